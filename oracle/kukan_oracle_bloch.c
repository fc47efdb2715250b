/*
 * kukan_oracle_bloch.c — FP64 CPU restatement of the complex (Bloch, k-point) path of the
 * RSPACE kukan v8f Kohn-Sham solver: nrc = 1, collinear spin (ncol = 1), no spin-orbit
 * (nso = 0).
 *
 * TEST INFRASTRUCTURE ONLY (the checker), like kukan_oracle.c: loaded by the tests, never by
 * the product library.  Every routine follows the reference routine cited above it (file:line
 * relative to RSPACEes.v08.01/kukan/) on one MPI rank, with its loop and summation order.
 * Complex arrays are Fortran complex*16: interleaved (re, im) doubles.  Products of complex
 * numbers are written out as the Fortran compiler evaluates (a, b)(c, d) = (ac - bd, ad + bc).
 *
 * Pins (tests/test_oracle_bloch.py): the Bloch plane-wave closed form (e^{i q.r} on the
 * periodic grid is an eigenvector of H_k with eigenvalue lambda(k + q)), the gauge identity
 * H_k u = e^{-i k.r} H (e^{i k.r} u) for k a reciprocal-lattice vector, a dense Hermitian H_k
 * (converged eigenvalues of kscg_c / rayleighritz_c), k = 0 against the real path, and the
 * compiled reference routines (tests/test_reference_bloch.py).
 */
#include "kukan_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int orc_set_error(int code, const char *msg);  /* kukan_oracle.c */
#define fail orc_set_error

static void *zalloc(size_t n) {
  void *p = calloc(n ? n : 1, 2 * sizeof(double));
  if (!p) {
    fprintf(stderr, "kukan_oracle_bloch: out of memory (%zu complex)\n", n);
    abort();
  }
  return p;
}

/* complex helpers on (re, im) pairs */
typedef struct {
  double r, i;
} zd;
static inline zd zmul(zd a, zd b) { return (zd){a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
static inline zd zadd(zd a, zd b) { return (zd){a.r + b.r, a.i + b.i}; }
static inline zd zsub(zd a, zd b) { return (zd){a.r - b.r, a.i - b.i}; }
static inline zd zconj(zd a) { return (zd){a.r, -a.i}; }
static inline zd zscal(double s, zd a) { return (zd){s * a.r, s * a.i}; }
#define ZG(p, k) (((zd *)(p))[k])

#define HX(ix, iy, iz) \
  ((size_t)((ix) + nf1) + (size_t)ex * ((size_t)((iy) + nf1) + (size_t)ey * (size_t)((iz) + nf1)))
#define IX3(ix, iy, iz) ((size_t)((ix)-1) + (size_t)ncpx * ((size_t)((iy)-1) + (size_t)ncpy * (size_t)((iz)-1)))

/* ------------------------------------------------------------------ halo
 * overlap_finitedifference8f.F90:327-645 overlap_finitedifference_c + :830-1010
 * overlap_fdcheck_c on one rank: the periodic wrap of each of the ncol components. */
int orc_overlap_finitedifference_c(int nperi, int na, int nb, int nc, int nf, int nf1, int nsnd,
                                   int ncol, double *v_cc) {
  const long ex = na + nf + nf1 + 1, ey = nb + nf + nf1 + 1, ez = nc + nf + nf1 + 1;
  for (int is = 0; is < ncol; is++) {
    zd *v = (zd *)v_cc + (size_t)is * ex * ey * ez;
    if (nperi >= 1)
      for (int ic = 1; ic <= nc; ic++)
        for (int ib = 1; ib <= nb; ib++)
          for (int iov = 1; iov <= nsnd; iov++) {
            zd lo = v[HX(na - nsnd + iov, ib, ic)], hi = v[HX(iov, ib, ic)];
            v[HX(-nsnd + iov, ib, ic)] = lo;
            v[HX(na + iov, ib, ic)] = hi;
          }
    if (nperi >= 2)
      for (int ic = 1; ic <= nc; ic++)
        for (int iov = 1; iov <= nsnd; iov++)
          for (int ia = 1; ia <= na; ia++) {
            zd lo = v[HX(ia, nb - nsnd + iov, ic)], hi = v[HX(ia, iov, ic)];
            v[HX(ia, -nsnd + iov, ic)] = lo;
            v[HX(ia, nb + iov, ic)] = hi;
          }
    if (nperi >= 3)
      for (int iov = 1; iov <= nsnd; iov++)
        for (int ib = 1; ib <= nb; ib++)
          for (int ia = 1; ia <= na; ia++) {
            zd lo = v[HX(ia, ib, nc - nsnd + iov)], hi = v[HX(ia, ib, iov)];
            v[HX(ia, ib, -nsnd + iov)] = lo;
            v[HX(ia, ib, nc + iov)] = hi;
          }
  }
  return 0;
}

/* ------------------------------------------------------------------ kslaplacian8f.F90:560-657
 * H_k u: the halo'd input is multiplied in place by e^{i k.r} (r from the unwrapped grid
 * index, x = jx dx - xmax - dx/2), kslaplacian_c_01 (:660-1226, the real three-sweep stencil
 * on each component: x neighbours with the centre and veff, then y, then z) writes workc,
 * and avcm(:,l) += e^{-i k.r} workc. */
int orc_kslaplacian_c(int ncpx, int ncpy, int ncpz, int neigmx, int ncol, int nvef, int nf,
                      double dx, double dy, double dz, double xmax, double ymax, double zmax,
                      double skpx, double skpy, double skpz, double *vcm, const double *veff,
                      int l, double *avcm, double *workc) {
  (void)neigmx; (void)nvef;
  if (ncol != 1) return fail(5, "kslaplacian_c: noncollinear spin (ncol = 2) is not part of this build");
  if (nf < 1 || nf > 8) return fail(1, "error in kslaplacian_c! This subroutine is for nf in 1-8.");
  const int nf1 = nf - 1;
  const long ex = ncpx + 2 * nf, ey = ncpy + 2 * nf;
  const size_t n = (size_t)ncpx * ncpy * ncpz;
  zd *v = (zd *)vcm, *w = (zd *)workc, *out = (zd *)avcm + (size_t)(l - 1) * n;
  for (int iz = -nf1; iz <= ncpz + nf; iz++)
    for (int iy = -nf1; iy <= ncpy + nf; iy++)
      for (int ix = -nf1; ix <= ncpx + nf; ix++) {
        const double x = ix * dx - xmax - 0.5 * dx, y = iy * dy - ymax - 0.5 * dy,
                     z = iz * dz - zmax - 0.5 * dz;
        const double th = skpx * x + skpy * y + skpz * z;
        v[HX(ix, iy, iz)] = zmul((zd){cos(th), sin(th)}, v[HX(ix, iy, iz)]);
      }
  double acfd[9], xx[9], yy[9], zz[9];
  orc_tools_definefdcoef(nf, acfd);
  for (int m = 0; m < 9; m++) {
    xx[m] = acfd[m] / dx / dx * (-0.5);
    yy[m] = acfd[m] / dy / dy * (-0.5);
    zz[m] = acfd[m] / dz / dz * (-0.5);
  }
  const double c0 = xx[0] + yy[0] + zz[0];
  memset(w, 0, sizeof(zd) * n);
  for (int iz = 1; iz <= ncpz; iz++)
    for (int iy = 1; iy <= ncpy; iy++)
      for (int ix = 1; ix <= ncpx; ix++) {
        zd s = w[IX3(ix, iy, iz)];
        for (int m = nf; m >= 1; m--) s = zadd(s, zscal(xx[m], v[HX(ix - m, iy, iz)]));
        s = zadd(s, zscal(c0 + veff[IX3(ix, iy, iz)], v[HX(ix, iy, iz)]));
        for (int m = 1; m <= nf; m++) s = zadd(s, zscal(xx[m], v[HX(ix + m, iy, iz)]));
        w[IX3(ix, iy, iz)] = s;
      }
  for (int iz = 1; iz <= ncpz; iz++)
    for (int iy = 1; iy <= ncpy; iy++)
      for (int ix = 1; ix <= ncpx; ix++) {
        zd s = w[IX3(ix, iy, iz)];
        for (int m = nf; m >= 1; m--) s = zadd(s, zscal(yy[m], v[HX(ix, iy - m, iz)]));
        for (int m = 1; m <= nf; m++) s = zadd(s, zscal(yy[m], v[HX(ix, iy + m, iz)]));
        w[IX3(ix, iy, iz)] = s;
      }
  for (int iz = 1; iz <= ncpz; iz++)
    for (int iy = 1; iy <= ncpy; iy++)
      for (int ix = 1; ix <= ncpx; ix++) {
        zd s = w[IX3(ix, iy, iz)];
        for (int m = nf; m >= 1; m--) s = zadd(s, zscal(zz[m], v[HX(ix, iy, iz - m)]));
        for (int m = 1; m <= nf; m++) s = zadd(s, zscal(zz[m], v[HX(ix, iy, iz + m)]));
        w[IX3(ix, iy, iz)] = s;
      }
  for (int iz = 1; iz <= ncpz; iz++)
    for (int iy = 1; iy <= ncpy; iy++)
      for (int ix = 1; ix <= ncpx; ix++) {
        const double x = ix * dx - xmax - 0.5 * dx, y = iy * dy - ymax - 0.5 * dy,
                     z = iz * dz - zmax - 0.5 * dz;
        const double th = skpx * x + skpy * y + skpz * z;
        out[IX3(ix, iy, iz)] = zadd(out[IX3(ix, iy, iz)], zmul((zd){cos(th), -sin(th)}, w[IX3(ix, iy, iz)]));
      }
  return 0;
}

/* ------------------------------------------------------------------ ksprecondition8f.F90:42-75 */
int orc_ksprecondition_c(int ncpx, int ncpy, int ncpz, int nf, int ncol, const double *vcm,
                         double *avcm) {
  const int nf1 = nf - 1;
  const long ex = ncpx + 2 * nf, ey = ncpy + 2 * nf;
  const double w = 1.0 / 12.0;
  const size_t hs = (size_t)ex * ey * (ncpz + 2 * nf), n = (size_t)ncpx * ncpy * ncpz;
  for (int is = 0; is < ncol; is++) {
    const zd *v = (const zd *)vcm + is * hs;
    zd *o = (zd *)avcm + is * n;
    for (int iz = 1; iz <= ncpz; iz++)
      for (int iy = 1; iy <= ncpy; iy++)
        for (int ix = 1; ix <= ncpx; ix++) {
          zd s = zscal(w, v[HX(ix - 1, iy, iz)]);
          s = zadd(s, zscal(w, v[HX(ix, iy - 1, iz)]));
          s = zadd(s, zscal(w, v[HX(ix, iy, iz - 1)]));
          s = zadd(s, zscal(0.5, v[HX(ix, iy, iz)]));
          s = zadd(s, zscal(w, v[HX(ix, iy, iz + 1)]));
          s = zadd(s, zscal(w, v[HX(ix, iy + 1, iz)]));
          s = zadd(s, zscal(w, v[HX(ix + 1, iy, iz)]));
          o[IX3(ix, iy, iz)] = s;
        }
  }
  return 0;
}

/* phase of list entry i of atom slot iaps: k.(lst + nat) h - k.h/2 per axis
 * (nonlocaloperation8f.F90:95-107, :343-356) */
static double list_phase(int i, int iaps, int na, int num_list, const int *lstx, const int *lsty,
                         const int *lstz, const int *natx, const int *naty, const int *natz,
                         double dx, double dy, double dz, double skpx, double skpy, double skpz) {
  const size_t k = (size_t)(i - 1) + (size_t)num_list * (iaps - 1);
  const int ix = lstx[k] + natx[na - 1], iy = lsty[k] + naty[na - 1], iz = lstz[k] + natz[na - 1];
  const double xx = ix * dx - 0.5 * dx, yy = iy * dy - 0.5 * dy, zz = iz * dz - 0.5 * dz;
  return skpx * xx + skpy * yy + skpz * zz;
}

/* ------------------------------------------------------------------ nonlocaloperation8f.F90:56-146
 * cpsep(j,a) += sum_i vnlocp(i,j,a) e^{i k.r_i} vcm2(lstvec2(i,a), l) */
int orc_nonlocaloperation_c_01(int natom, int nprjmx, int num_spe, int num_ppcell, int num_list,
                               int neigmx, int ncol, int l, int ncpx, int ncpy, int ncpz,
                               int key_natpri_in, int key_natpri_inps, double dx, double dy,
                               double dz, double skpx, double skpy, double skpz, const int *nprj,
                               const int *indspe, const int *natinf, const int *natpri,
                               const int *naps, const int *lstvec2, const int *lstx,
                               const int *lsty, const int *lstz, const int *natx, const int *naty,
                               const int *natz, const double *vnlocp, const double *vcm2,
                               double *cpsep, double *vcccm) {
  (void)num_spe; (void)num_ppcell; (void)neigmx;
  const size_t n = (size_t)ncpx * ncpy * ncpz;
  zd tmp[64];
  for (int na = 1; na <= natom; na++) {
    if (natpri[na - 1] != key_natpri_in && natpri[na - 1] != key_natpri_inps) continue;
    const int iaps = naps[na - 1], iatinf = natinf[na - 1], iprj = nprj[indspe[na - 1] - 1];
    for (int j = 0; j < iprj * ncol; j++) tmp[j] = (zd){0.0, 0.0};
    for (int ns = 1; ns <= ncol; ns++) {
      const zd *v = (const zd *)vcm2 + n * ((size_t)(l - 1) + (size_t)neigmx * (ns - 1));
      zd *vc = (zd *)vcccm + (size_t)num_list * (ns - 1);
      for (int i = 1; i <= iatinf; i++) {
        const double th = list_phase(i, iaps, na, num_list, lstx, lsty, lstz, natx, naty, natz, dx,
                                     dy, dz, skpx, skpy, skpz);
        const int ii = lstvec2[(size_t)(i - 1) + (size_t)num_list * (iaps - 1)];
        vc[i - 1] = zmul((zd){cos(th), sin(th)}, v[ii - 1]);
      }
    }
    for (int ns = 1; ns <= ncol; ns++) {
      const zd *vc = (const zd *)vcccm + (size_t)num_list * (ns - 1);
      for (int j = 1; j <= iprj; j++) {
        const double *p = vnlocp + (size_t)num_list * ((j - 1) + (size_t)nprjmx * (iaps - 1));
        zd t = tmp[(j - 1) + iprj * (ns - 1)];
        for (int i = 0; i < iatinf; i++) t = zadd(t, zscal(p[i], vc[i]));
        tmp[(j - 1) + iprj * (ns - 1)] = t;
      }
    }
    for (int ns = 1; ns <= ncol; ns++)
      for (int j = 1; j <= iprj; j++) {
        zd *c = (zd *)cpsep + (size_t)(j - 1) + (size_t)nprjmx * ((iaps - 1) + (size_t)num_ppcell * (ns - 1));
        *c = zadd(*c, tmp[(j - 1) + iprj * (ns - 1)]);
      }
  }
  return 0;
}

/* ------------------------------------------------------------------ nonlocaloperation8f.F90:238-393
 * nhs = 1: c = D c_in / dV (dijsss(:,:,ns1,atom)); nhs = 0: c = q c_in / dV (species);
 * avcm(lstvec2(i), l) += e^{-i k.r_i} sum_j c_j vnlocp(i,j).  Collinear (ncol = 1), no
 * spin-orbit term (nso = 0). */
int orc_nonlocaloperation_c_02(int nhs, int nso, int natom, int num_spe, int neigmx, int nums,
                               int nprjmx, int ncol, int ns1, int l, int num_list, int num_ppcell,
                               int ncpx, int ncpy, int ncpz, int npxmax, int npymax, int npzmax,
                               int key_natpri_in, int key_natpri_inps, int key_soc_calc,
                               double dx, double dy, double dz, double skpx, double skpy,
                               double skpz, const int *nprj, const int *indspe, const int *natinf,
                               const int *lstvec2, const int *natpri, const int *naps,
                               const int *natsoc, const int *lstx, const int *lsty,
                               const int *lstz, const int *natx, const int *naty, const int *natz,
                               const double *vnlocp, const double *cpsep, const double *dijsss,
                               const double *dijsoc, double *avcm, double *vcccm, double *avc) {
  (void)num_spe; (void)npxmax; (void)npymax; (void)npzmax; (void)key_soc_calc; (void)natsoc;
  (void)dijsoc; (void)vcccm;
  if (nhs > 1 || nso > 1 || nhs < 0 || nso < 0) return fail(1, "error occurs in nonlocaloperation_c_02");
  if (nhs == 0 && ns1 != 1) return fail(1, "nonlocaloperation_c_02: nhs=0 and ns1/=1");
  if (ncol != 1 || nso != 0)
    return fail(5, "nonlocaloperation_c_02: noncollinear spin and spin-orbit are not part of this build");
  const size_t n = (size_t)ncpx * ncpy * ncpz;
  const int d3 = nums * ncol * nhs - nhs + 1;
  const double dxyzin = 1.0 / (dx * dy * dz);
  zd ct[64], ct2[64];
  zd *out = (zd *)avcm + n * (size_t)(l - 1);
  zd *a = (zd *)avc;
  (void)neigmx;
  for (int na = 1; na <= natom; na++) {
    if (natpri[na - 1] != key_natpri_in && natpri[na - 1] != key_natpri_inps) continue;
    const int iaps = naps[na - 1], iatinf = natinf[na - 1], ispe = indspe[na - 1];
    const int iprj = nprj[ispe - 1];
    for (int i = 1; i <= iprj; i++) {
      ct2[i - 1] = (zd){0.0, 0.0};
      ct[i - 1] = zscal(dxyzin, ZG(cpsep, (size_t)(i - 1) + (size_t)nprjmx * (iaps - 1)));
    }
    const int i3 = nhs == 1 ? ns1 : 1, i4 = nhs == 1 ? na : ispe;
    const double *X = dijsss + (size_t)nprjmx * nprjmx * ((i3 - 1) + (size_t)d3 * (i4 - 1));
    for (int j = 1; j <= iprj; j++)
      for (int i = 1; i <= iprj; i++)
        ct2[i - 1] = zadd(ct2[i - 1], zscal(X[(i - 1) + (size_t)nprjmx * (j - 1)], ct[j - 1]));
    const double *p = vnlocp + (size_t)num_list * nprjmx * (iaps - 1);
    for (int i = 0; i < iatinf; i++) a[i] = zscal(p[i], ct2[0]);
    for (int j = 2; j <= iprj; j++)
      for (int i = 0; i < iatinf; i++) a[i] = zadd(a[i], zscal(p[i + (size_t)num_list * (j - 1)], ct2[j - 1]));
    for (int i = 1; i <= iatinf; i++) {
      const double th = list_phase(i, iaps, na, num_list, lstx, lsty, lstz, natx, naty, natz, dx, dy,
                                   dz, skpx, skpy, skpz);
      const int j = lstvec2[(size_t)(i - 1) + (size_t)num_list * (iaps - 1)];
      out[j - 1] = zadd(out[j - 1], zmul((zd){cos(th), -sin(th)}, a[i - 1]));
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ shared context */
typedef struct {
  int ncpx, ncpy, ncpz, nf, nperi;
  double dx, dy, dz, dv, xmax, ymax, zmax, kx, ky, kz;
  int natom, num_spe, nums, nprjmx, num_list, num_ppcell;
  int key_in, key_inps;
  const int *nprj, *indspe, *natinf, *lstvec2, *natpri, *naps, *lstx, *lsty, *lstz, *natx, *naty, *natz;
  const double *vnlocp, *dij, *sss, *veff;
  long n, nh, n1;
  zd *vh, *work, *vc, *avc;
} zctx_t;

static void zctx_alloc(zctx_t *c) {
  c->n = (long)c->ncpx * c->ncpy * c->ncpz;
  c->nh = (long)(c->ncpx + 2 * c->nf) * (c->ncpy + 2 * c->nf) * (c->ncpz + 2 * c->nf);
  c->n1 = (long)c->nprjmx * c->num_ppcell;
  c->dv = c->dx * c->dy * c->dz;
  c->vh = (zd *)zalloc(c->nh);
  c->work = (zd *)zalloc(c->n);
  c->vc = (zd *)zalloc(c->num_list);
  c->avc = (zd *)zalloc(c->num_list);
}
static void zctx_free(zctx_t *c) {
  free(c->vh); free(c->work); free(c->vc); free(c->avc);
}

static void zload_halo(zctx_t *c, const zd *u, int nsnd) {
  const int nf = c->nf, nf1 = nf - 1, ncpx = c->ncpx, ncpy = c->ncpy, ncpz = c->ncpz;
  const long ex = ncpx + 2 * nf, ey = ncpy + 2 * nf;
  memset(c->vh, 0, sizeof(zd) * c->nh);
  for (int iz = 1; iz <= ncpz; iz++)
    for (int iy = 1; iy <= ncpy; iy++)
      for (int ix = 1; ix <= ncpx; ix++) c->vh[HX(ix, iy, iz)] = u[IX3(ix, iy, iz)];
  orc_overlap_finitedifference_c(c->nperi, ncpx, ncpy, ncpz, nf, nf1, nsnd, 1, (double *)c->vh);
}

/* out = H_k u with given <p|u> (kscg8f.F90:646-697): zero, nonlocal c_02 (nhs = 1), then
 * kslaplacian_c on the halo'd u. */
static int zapply_h(zctx_t *c, const zd *u, const zd *cp, zd *out) {
  memset(out, 0, sizeof(zd) * c->n);
  zload_halo(c, u, c->nf);
  int rc = orc_nonlocaloperation_c_02(
      1, 0, c->natom, c->num_spe, 1, c->nums, c->nprjmx, 1, 1, 1, c->num_list, c->num_ppcell,
      c->ncpx, c->ncpy, c->ncpz, 0, 0, 0, c->key_in, c->key_inps, 0, c->dx, c->dy, c->dz, c->kx,
      c->ky, c->kz, c->nprj, c->indspe, c->natinf, c->lstvec2, c->natpri, c->naps, NULL, c->lstx,
      c->lsty, c->lstz, c->natx, c->naty, c->natz, c->vnlocp, (const double *)cp, c->dij, NULL,
      (double *)out, (double *)c->vc, (double *)c->avc);
  if (rc) return rc;
  return orc_kslaplacian_c(c->ncpx, c->ncpy, c->ncpz, 1, 1, 1, c->nf, c->dx, c->dy, c->dz, c->xmax,
                           c->ymax, c->zmax, c->kx, c->ky, c->kz, (double *)c->vh, c->veff, 1,
                           (double *)out, (double *)c->work);
}

/* cp = <p|u> (nonlocaloperation_c_01 on zeroed cp) */
static void zproj(zctx_t *c, const zd *u, zd *cp) {
  memset(cp, 0, sizeof(zd) * c->n1);
  orc_nonlocaloperation_c_01(c->natom, c->nprjmx, c->num_spe, c->num_ppcell, c->num_list, 1, 1, 1,
                             c->ncpx, c->ncpy, c->ncpz, c->key_in, c->key_inps, c->dx, c->dy,
                             c->dz, c->kx, c->ky, c->kz, c->nprj, c->indspe, c->natinf, c->natpri,
                             c->naps, c->lstvec2, c->lstx, c->lsty, c->lstz, c->natx, c->naty,
                             c->natz, c->vnlocp, (const double *)u, (double *)cp, (double *)c->vc);
}

/* su = u + S-part with given <p|u> (nonlocaloperation_c_02, nhs = 0, sss) */
static void zapply_s(zctx_t *c, const zd *u, const zd *cp, zd *su) {
  memcpy(su, u, sizeof(zd) * c->n);
  orc_nonlocaloperation_c_02(0, 0, c->natom, c->num_spe, 1, c->nums, c->nprjmx, 1, 1, 1, c->num_list,
                             c->num_ppcell, c->ncpx, c->ncpy, c->ncpz, 0, 0, 0, c->key_in,
                             c->key_inps, 0, c->dx, c->dy, c->dz, c->kx, c->ky, c->kz, c->nprj,
                             c->indspe, c->natinf, c->lstvec2, c->natpri, c->naps, NULL, c->lstx,
                             c->lsty, c->lstz, c->natx, c->naty, c->natz, c->vnlocp,
                             (const double *)cp, c->sss, NULL, (double *)su, (double *)c->vc,
                             (double *)c->avc);
}

/* out = K r (kscg8f.F90:720-748) */
static void zprecond(zctx_t *c, const zd *r, zd *out) {
  zload_halo(c, r, 1);
  orc_ksprecondition_c(c->ncpx, c->ncpy, c->ncpz, c->nf, 1, (const double *)c->vh, (double *)out);
}

/* orthogonalization_c_04 (:796-840): sum_i conj(a_i) b_i */
static zd zdotc(long n, const zd *a, const zd *b) {
  zd s = {0.0, 0.0};
  for (long i = 0; i < n; i++) s = zadd(s, zmul(zconj(a[i]), b[i]));
  return s;
}
/* orthogonalization_c_05 (:843-877): sum re a re b + im a im b */
static double zdotr(long n, const zd *a, const zd *b) {
  double s = 0.0;
  for (long i = 0; i < n; i++) s = s + a[i].r * b[i].r + a[i].i * b[i].i;
  return s;
}
/* orthogonalization_c_03 (:747-793): v -= sum_j ca(j) a(:,j) */
static void zr03(long n, int lmax, zd *v, const zd *a, const zd *ca) {
  for (int j = 0; j < lmax; j++) {
    const zd *aj = a + (size_t)n * j;
    for (long i = 0; i < n; i++) v[i] = zsub(v[i], zmul(ca[j], aj[i]));
  }
}
/* multc (:915-932): a0(:,j) -= sum_k b(j,k) a1(:,k) */
static void zmultc(long n, int m, int l, zd *a0, const zd *a1, const zd *b) {
  for (int k = 0; k < l; k++)
    for (int j = 0; j < m; j++)
      for (long i = 0; i < n; i++)
        a0[i + n * j] = zsub(a0[i + n * j], zmul(b[j + (size_t)m * k], a1[i + n * k]));
}
static int znormalise(zctx_t *c, zd *v, zd *sv, zd *cp, const char *msg) {
  const double vecnor = zdotr(c->n, v, sv) * c->dv;
  if (vecnor <= 0.0) return fail(2, msg);
  const double t = 1.0 / sqrt(vecnor);
  for (long i = 0; i < c->n; i++) v[i] = zscal(t, v[i]);
  for (long i = 0; i < c->n; i++) sv[i] = zscal(t, sv[i]);
  for (long i = 0; i < c->n1; i++) cp[i] = zscal(t, cp[i]);
  return 0;
}

/* ------------------------------------------------------------------ orthogonalization8f.F90:379-724 */
static int zortho_c_01(zctx_t *c, int neigmx, int l_inp, int nflag, int key_cmpt, zd *psi, zd *spsi,
                       zd *cps) {
  const long n = c->n, n1 = c->n1;
  int mstep = 32;
  if (neigmx > 1024) mstep = 64;
  if (neigmx > 2048) mstep = 128;
  if (neigmx > 4096) mstep = 256;
  const int l_stt = l_inp == 0 ? 1 : l_inp, l_end = l_inp == 0 ? neigmx : l_inp;
#define PSI(l) (psi + (size_t)n * ((l)-1))
#define SPSI(l) (spsi + (size_t)n * ((l)-1))
#define CPS(l) (cps + (size_t)n1 * ((l)-1))
  if (nflag == key_cmpt) {
    for (int l = l_stt; l <= l_end; l++) zproj(c, PSI(l), CPS(l));
    for (int l = l_stt; l <= l_end; l++) zapply_s(c, PSI(l), CPS(l), SPSI(l));
  }
  zd *av = (zd *)zalloc(neigmx), *cm = (zd *)zalloc((size_t)mstep * mstep);
  int rc = 0, l = l_stt;
  while (l < l_end + 1) {
    if (l > mstep && l_end - l + 1 >= mstep) {
      for (int j = 1; j <= l - 1; j += mstep) {
        /* zgemm('c','n'): cm(a,b) = sum conj(psi_{l+a}) spsi_{j+b}; then conj * dV */
        for (int a = 0; a < mstep; a++)
          for (int b = 0; b < mstep; b++)
            cm[a + mstep * b] = zscal(c->dv, zconj(zdotc(n, PSI(l + a), SPSI(j + b))));
        /* zgemm('n','t'): psi_{l+a} -= sum_b psi_{j+b} cm(a,b) */
        for (int a = 0; a < mstep; a++) {
          zd *va = PSI(l + a), *sa = SPSI(l + a);
          for (int b = 0; b < mstep; b++) {
            const zd cab = cm[a + mstep * b];
            const zd *vb = PSI(j + b), *sb = SPSI(j + b);
            for (long i = 0; i < n; i++) {
              va[i] = zsub(va[i], zmul(vb[i], cab));
              sa[i] = zsub(sa[i], zmul(sb[i], cab));
            }
          }
        }
        zmultc(n1, mstep, mstep, CPS(l), CPS(j), cm);
      }
      if ((rc = znormalise(c, PSI(l), SPSI(l), CPS(l), "orthogonalization_c_01 (1): wf has negative norm"))) break;
      for (int l0 = l + 1; l0 <= l + mstep - 1; l0++) {
        for (int i = 0; i < l0 - l; i++) av[i] = zscal(c->dv, zdotc(n, SPSI(l + i), PSI(l0)));
        zr03(n, l0 - l, PSI(l0), PSI(l), av);
        zr03(n, l0 - l, SPSI(l0), SPSI(l), av);
        zmultc(n1, 1, l0 - l, CPS(l0), CPS(l), av);
        if ((rc = znormalise(c, PSI(l0), SPSI(l0), CPS(l0), "orthogonalization_c_01 (2): wf has negative norm"))) break;
      }
      if (rc) break;
      l += mstep;
    } else {
      for (int j = 1; j <= l - 1; j++) av[j - 1] = zscal(c->dv, zdotc(n, SPSI(j), PSI(l)));
      zr03(n, l - 1, PSI(l), PSI(1), av);
      zr03(n, l - 1, SPSI(l), SPSI(1), av);
      zmultc(n1, 1, l - 1, CPS(l), CPS(1), av);
      if ((rc = znormalise(c, PSI(l), SPSI(l), CPS(l), "orthogonalization_c_01 (3): wf has negative norm"))) break;
      l += 1;
    }
  }
  free(av);
  free(cm);
  return rc;
#undef PSI
#undef SPSI
#undef CPS
}

static void zctx_set(zctx_t *c, int ncpx, int ncpy, int ncpz, int nf, int nperi, double dx, double dy,
                     double dz, int natom, int num_spe, int nums, int nprjmx, int num_list,
                     int num_ppcell, int key_in, int key_inps, const int *nprj, const int *indspe,
                     const int *natinf, const int *lstvec2, const int *natpri, const int *naps,
                     const int *lstx, const int *lsty, const int *lstz, const int *natx,
                     const int *naty, const int *natz, const double *vnlocp, const double *sss,
                     double skpx, double skpy, double skpz) {
  memset(c, 0, sizeof *c);
  c->ncpx = ncpx; c->ncpy = ncpy; c->ncpz = ncpz; c->nf = nf; c->nperi = nperi;
  c->dx = dx; c->dy = dy; c->dz = dz;
  c->xmax = 0.5 * ncpx * dx; c->ymax = 0.5 * ncpy * dy; c->zmax = 0.5 * ncpz * dz;
  c->kx = skpx; c->ky = skpy; c->kz = skpz;
  c->natom = natom; c->num_spe = num_spe; c->nums = nums; c->nprjmx = nprjmx;
  c->num_list = num_list; c->num_ppcell = num_ppcell; c->key_in = key_in; c->key_inps = key_inps;
  c->nprj = nprj; c->indspe = indspe; c->natinf = natinf; c->lstvec2 = lstvec2;
  c->natpri = natpri; c->naps = naps; c->lstx = lstx; c->lsty = lsty; c->lstz = lstz;
  c->natx = natx; c->naty = naty; c->natz = natz; c->vnlocp = vnlocp; c->sss = sss;
}

int orc_orthogonalization_c_01(int natom, int neigmx, int nprjmx, int nums, int ncol, int numk,
                               int nk, int ns1, int l_inp, int num_list, int num_ppcell,
                               int num_spe, int nflag, int ndisp, int ncpx, int ncpy, int ncpz,
                               int npxmax, int npymax, int npzmax, int key_ortho_cmpt_innerproduct,
                               int key_natpri_in, int key_natpri_inps, double dx, double dy,
                               double dz, double skpx, double skpy, double skpz, const int *nprj,
                               const int *indspe, const int *natpri, const int *naps,
                               const int *natinf, const int *lstvec2, const int *latom,
                               const int *lstx, const int *lsty, const int *lstz, const int *natx,
                               const int *naty, const int *natz, const double *vnlocp,
                               const double *sss, double *sveccm, double *ssvcm, double *cspsep) {
  (void)numk; (void)ndisp; (void)npxmax; (void)npymax; (void)npzmax; (void)latom;
  if (ncol != 1) return fail(5, "orthogonalization_c_01: noncollinear spin (ncol = 2) is not part of this build");
  zctx_t c;
  zctx_set(&c, ncpx, ncpy, ncpz, 1, 0, dx, dy, dz, natom, num_spe, nums, nprjmx, num_list, num_ppcell,
           key_natpri_in, key_natpri_inps, nprj, indspe, natinf, lstvec2, natpri, naps, lstx, lsty,
           lstz, natx, naty, natz, vnlocp, sss, skpx, skpy, skpz);
  zctx_alloc(&c);
  const size_t off = (size_t)c.n * neigmx * ((ns1 - 1) + (size_t)nums * (nk - 1));
  const size_t off1 = (size_t)c.n1 * neigmx * ((ns1 - 1) + (size_t)nums * (nk - 1));
  int rc = zortho_c_01(&c, neigmx, l_inp, nflag, key_ortho_cmpt_innerproduct, (zd *)sveccm + off,
                       (zd *)ssvcm + off, (zd *)cspsep + off1);
  zctx_free(&c);
  return rc;
}

/* orthogonalization8f.F90:47-90 */
int orc_orthogonalization_c(int natom, int neigmx, int nprjmx, int nums, int ncol, int numk,
                            int num_ppcell, int num_spe, int northo, int num_list, int nflag,
                            int ndisp, int ncpx, int ncpy, int ncpz, int npxmax, int npymax,
                            int npzmax, int key_ortho_cmpt_innerproduct, int key_natpri_in,
                            int key_natpri_inps, double dx, double dy, double dz, const int *nprj,
                            const int *indspe, const int *natpri, const int *naps,
                            const int *natinf, const int *lstvec2, const int *latom,
                            const int *ntyppp, const int *lstx, const int *lsty, const int *lstz,
                            const int *natx, const int *naty, const int *natz,
                            const double *skpxx, const double *skpyy, const double *skpzz,
                            const double *vnlocp, const double *sss, double *sveccm,
                            double *ssvcm, double *cspsep) {
  (void)ntyppp;
  for (int nk = 1; nk <= numk; nk++)
    for (int ns = 1; ns <= nums + 1 - ncol; ns++)
      for (int loop = 1; loop <= northo; loop++) {
        int rc = orc_orthogonalization_c_01(
            natom, neigmx, nprjmx, nums, ncol, numk, nk, ns, 0, num_list, num_ppcell, num_spe, nflag,
            ndisp, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct,
            key_natpri_in, key_natpri_inps, dx, dy, dz, skpxx[nk - 1], skpyy[nk - 1], skpzz[nk - 1],
            nprj, indspe, natpri, naps, natinf, lstvec2, latom, lstx, lsty, lstz, natx, naty, natz,
            vnlocp, sss, sveccm, ssvcm, cspsep);
        if (rc) return rc;
      }
  return 0;
}

/* ------------------------------------------------------------------ kscg8f.F90:492-1032, 1053-1274 */
int orc_kscg_c(int nso, int natom, int num_spe, int neigmx, int nums, int ncol, int nspv, int numk,
               int nperi, int northo, int ns1, int nk, int nf, int nkscg, int nsdmax,
               int nprecon_cg, int nprjmx, int num_list, int num_ppcell, int ndisp, int ncpx,
               int ncpy, int ncpz, int npxmax, int npymax, int npzmax,
               int key_ortho_nocmpt_innerproduct, int key_ortho_cmpt_innerproduct,
               int key_natpri_in, int key_natpri_inps, int key_natpri_out, int key_soc_calc,
               double dx, double dy, double dz, double skpx, double skpy, double skpz,
               double epssd, const int *indspe, const int *natpri, const int *naps,
               const int *nprj, const int *natinf, const int *lstvec2, const int *latom,
               const int *natsoc, const int *ntyppp, const int *lstx, const int *lsty,
               const int *lstz, const int *natx, const int *naty, const int *natz,
               const double *dij, const double *dijsoc, const double *veff, const double *vnlocp,
               const double *sss, int *ksconv, int *ksitmax, double *sval, double *cspsep,
               double *sveccm, double *ssvcm, double *residual_states, double *hsvcm) {
  (void)numk; (void)ndisp; (void)npxmax; (void)npymax; (void)npzmax; (void)key_natpri_out;
  (void)key_soc_calc; (void)latom; (void)natsoc; (void)ntyppp; (void)dijsoc;
  if (ncol != 1 || nso != 0)
    return fail(5, "kscg_c: noncollinear spin and spin-orbit are not part of this build");
  zctx_t c;
  zctx_set(&c, ncpx, ncpy, ncpz, nf, nperi, dx, dy, dz, natom, num_spe, nums, nprjmx, num_list,
           num_ppcell, key_natpri_in, key_natpri_inps, nprj, indspe, natinf, lstvec2, natpri, naps,
           lstx, lsty, lstz, natx, naty, natz, vnlocp, sss, skpx, skpy, skpz);
  const int nsv = ns1 < nspv ? ns1 : nspv;
  c.dij = dij;
  zctx_alloc(&c);
  c.veff = veff + (size_t)(nsv - 1) * c.n;
  const long n = c.n, n1 = c.n1;
  const size_t off = (size_t)n * neigmx * ((ns1 - 1) + (size_t)nums * (nk - 1));
  const size_t off1 = (size_t)n1 * neigmx * ((ns1 - 1) + (size_t)nums * (nk - 1));
  zd *Psi = (zd *)sveccm + off, *SPsi = (zd *)ssvcm + off, *Cps = (zd *)cspsep + off1;
  double *sv = sval + (size_t)neigmx * ((ns1 - 1) + (size_t)nums * (nk - 1));
  zd *rvec = (zd *)zalloc(n), *pvec = (zd *)zalloc(n), *pv = (zd *)zalloc(n), *rv = (zd *)zalloc(n);
  zd *spv = (zd *)zalloc(n), *av = (zd *)zalloc(n), *asv = (zd *)zalloc(n);
  zd *avc = (zd *)zalloc(neigmx), *crp = (zd *)zalloc(n1), *cpp = (zd *)zalloc(n1), *cp = (zd *)zalloc(n1);
  const int nsdmax1 = 10, nsdmax2 = 200;
  int rc = 0;
  for (int l = 1; l <= neigmx && !rc; l++) {
    zd *psi = Psi + (size_t)n * (l - 1), *spsi = SPsi + (size_t)n * (l - 1);
    zd *hpsi = (zd *)hsvcm + (size_t)n * (l - 1), *cps = Cps + (size_t)n1 * (l - 1);
    const int key = northo == 2 ? key_ortho_cmpt_innerproduct : key_ortho_nocmpt_innerproduct;
    for (int loop = 1; loop <= northo && !rc; loop++)
      rc = zortho_c_01(&c, neigmx, l, key, key_ortho_cmpt_innerproduct, Psi, SPsi, Cps);
    if (rc) break;
    memcpy(cp, cps, sizeof(zd) * n1);
    if ((rc = zapply_h(&c, psi, cp, hpsi))) break;
    int iiter = 0, l_converged = 0;
    zd r2r = {0.0, 0.0}, r1r;
    while (iiter < nsdmax && !l_converged) {
      iiter++;
      int doreset = (iiter == 1) || (iiter == nsdmax1);
      if (nsdmax2 > 0) doreset = doreset || ((iiter - 1) % nsdmax2 == 0);
      if (doreset) {
        r2r = (zd){1.0, 0.0};
        memset(pvec, 0, sizeof(zd) * n);
        memset(rvec, 0, sizeof(zd) * n);
        memset(pv, 0, sizeof(zd) * n);
        memset(spv, 0, sizeof(zd) * n);
        memset(cpp, 0, sizeof(zd) * n1);
      }
      /* eigenvalue and residual (:702-716) */
      r1r = r2r;
      const double valre = zdotr(n, psi, hpsi) * c.dv;
      sv[l - 1] = valre;
      for (long i = 0; i < n; i++) rvec[i] = zscal(-1.0, zsub(hpsi[i], zscal(valre, spsi[i])));
      /* precondition (:719-748) */
      if (nprecon_cg == 1) zprecond(&c, rvec, rv);
      else memcpy(rv, rvec, sizeof(zd) * n);
      /* K r orthogonal to psi_1..l in the S metric (:751-778) */
      for (int loop = 1; loop <= northo; loop++) {
        for (int j = 1; j <= l; j++) avc[j - 1] = zscal(c.dv, zdotc(n, SPsi + (size_t)n * (j - 1), rv));
        zr03(n, l, rv, Psi, avc);
      }
      /* <p|Kr>, S Kr (:779-822) */
      zproj(&c, rv, cp);
      memcpy(crp, cp, sizeof(zd) * n1);
      zapply_s(&c, rv, crp, asv);
      /* r2r = dV sum conj(K r) r (orthogonalization_c_04 with v1 = r, v2 = K r); residual
       * = |r2r| (:823-834) */
      r2r = zscal(c.dv, zdotc(n, rv, rvec));
      residual_states[l - 1] = hypot(r2r.r, r2r.i);
      /* conjugate direction (:837-861): kscg_c_03 with conj(r1r) r2r / |r1r|^2 */
      if (nkscg == 1) {
        const double a2 = hypot(r1r.r, r1r.i);
        const zd t = zscal(1.0 / (a2 * a2), zmul(zconj(r1r), r2r));
        for (long i = 0; i < n; i++) {
          pvec[i] = zadd(rv[i], zmul(t, pv[i]));
          pv[i] = pvec[i];
          spv[i] = zadd(asv[i], zmul(t, spv[i]));
        }
        /* cppsep = crpsep + r2r / r1r * cppsep (Fortran complex division) */
        const double den = r1r.r * r1r.r + r1r.i * r1r.i;
        const zd q = {(r2r.r * r1r.r + r2r.i * r1r.i) / den, (r2r.i * r1r.r - r2r.r * r1r.i) / den};
        for (long i = 0; i < n1; i++) cpp[i] = zadd(crp[i], zmul(q, cpp[i]));
      } else {
        memcpy(pvec, rv, sizeof(zd) * n);
        memcpy(spv, asv, sizeof(zd) * n);
        memcpy(cpp, crp, sizeof(zd) * n1);
      }
      /* p orthogonal to psi_1..l, normalised (:864-938) */
      memcpy(cp, cpp, sizeof(zd) * n1);
      memcpy(asv, spv, sizeof(zd) * n);
      for (int loop = 1; loop <= northo; loop++) {
        for (int j = 1; j <= l; j++) avc[j - 1] = zscal(c.dv, zdotc(n, SPsi + (size_t)n * (j - 1), pvec));
        zr03(n, l, pvec, Psi, avc);
        zr03(n, l, asv, SPsi, avc);
        zmultc(n1, 1, l, cp, Cps, avc);
      }
      {
        const double vecnor = zdotr(n, pvec, asv) * c.dv;
        if (vecnor <= 0.0) {
          rc = fail(2, "kscg_c: pvec has negative norm");
          break;
        }
        const double t = 1.0 / sqrt(vecnor);
        for (long i = 0; i < n; i++) pvec[i] = zscal(t, pvec[i]);
        for (long i = 0; i < n; i++) asv[i] = zscal(t, asv[i]);
        for (long i = 0; i < n1; i++) cp[i] = zscal(t, cp[i]);
      }
      /* H p (:941-982), pap and sap (kscg_c_04) */
      if ((rc = zapply_h(&c, pvec, cp, av))) break;
      const double pap = zdotr(n, pvec, av) * c.dv;
      const zd sap = zscal(c.dv, zdotc(n, psi, av));
      /* kscg_c_05 (:1233-1274) */
      {
        const double eps = 1.0e-14, s2 = sap.r * sap.r + sap.i * sap.i;
        const double sval_l = sv[l - 1];
        const double gamma = (sval_l + pap) * 0.5 - sqrt((sval_l - pap) * (sval_l - pap) * 0.25 + s2);
        double alph;
        zd cbeta;
        if (fabs((sval_l - gamma) / gamma) > eps) {
          const double d = sqrt(s2 + (sval_l - gamma) * (sval_l - gamma));
          const zd calph = {sap.r / d, sap.i / d};
          cbeta = (zd){-(sval_l - gamma) / d, 0.0};
          alph = hypot(calph.r, calph.i);
          if (alph > eps) {
            /* cbeta * alph / calph (Fortran complex division) */
            const zd num = zscal(alph, cbeta);
            const double dd = calph.r * calph.r + calph.i * calph.i;
            cbeta = (zd){(num.r * calph.r + num.i * calph.i) / dd, (num.i * calph.r - num.r * calph.i) / dd};
          }
        } else {
          alph = 1.0;
          cbeta = (zd){0.0, 0.0};
          l_converged = 1;
        }
        for (long i = 0; i < n; i++) {
          psi[i] = zadd(zscal(alph, psi[i]), zmul(cbeta, pvec[i]));
          hpsi[i] = zadd(zscal(alph, hpsi[i]), zmul(cbeta, av[i]));
          spsi[i] = zadd(zscal(alph, spsi[i]), zmul(cbeta, asv[i]));
        }
        for (long i = 0; i < n1; i++) cps[i] = zadd(zscal(alph, cps[i]), zmul(cbeta, cp[i]));
      }
      l_converged = sqrt(residual_states[l - 1]) < epssd;
      *ksconv = *ksconv && (l_converged || (iiter < nsdmax));
    }
    if (iiter > *ksitmax) *ksitmax = iiter;
  }
  free(rvec); free(pvec); free(pv); free(rv); free(spv); free(av); free(asv);
  free(avc); free(crp); free(cpp); free(cp);
  zctx_free(&c);
  return rc;
}

/* ------------------------------------------------------------------ Hermitian eigensolver
 * Cyclic Jacobi for a Hermitian matrix (the complex Jacobi rotation of Golub & Van Loan,
 * Matrix Computations, section 8.5, applied to the Hermitian case): a(n,n) column-major
 * (upper triangle used, as zheev('v','u')) is overwritten by the eigenvectors, w by the
 * eigenvalues in ascending order.  Each eigenvector is scaled so that its
 * largest-magnitude component (lowest index on ties) is real and positive. */
int orc_zheev(int n, double *a_, double *w) {
  zd *a = (zd *)a_;
#define A(i, j) a[(size_t)(i) + (size_t)n * (j)]
#define V(i, j) v[(size_t)(i) + (size_t)n * (j)]
  zd *v = (zd *)zalloc((size_t)n * n);
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < j; i++) A(j, i) = zconj(A(i, j)); /* from the upper triangle */
    A(j, j).i = 0.0;
    V(j, j) = (zd){1.0, 0.0};
  }
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const double m2 = A(i, j).r * A(i, j).r + A(i, j).i * A(i, j).i;
        tot += m2;
        if (i != j) off += m2;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const zd apq = A(p, q);
        const double g = hypot(apq.r, apq.i);
        if (g == 0.0) continue;
        /* rotation in the (p, q) plane: phase e = apq / |apq|, real Jacobi on (app, |apq|, aqq) */
        const zd e = {apq.r / g, apq.i / g};
        const double app = A(p, p).r, aqq = A(q, q).r;
        const double tau = (aqq - app) / (2.0 * g);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
        /* columns: [a_p a_q] <- [a_p a_q] [[c, s e], [-s conj(e), c]] */
        for (int k = 0; k < n; k++) {
          const zd akp = A(k, p), akq = A(k, q);
          A(k, p) = zsub(zscal(cs, akp), zscal(sn, zmul(zconj(e), akq)));
          A(k, q) = zadd(zscal(sn, zmul(e, akp)), zscal(cs, akq));
          const zd vkp = V(k, p), vkq = V(k, q);
          V(k, p) = zsub(zscal(cs, vkp), zscal(sn, zmul(zconj(e), vkq)));
          V(k, q) = zadd(zscal(sn, zmul(e, vkp)), zscal(cs, vkq));
        }
        /* rows: conjugate transpose of the same rotation */
        for (int k = 0; k < n; k++) {
          const zd apk = A(p, k), aqk = A(q, k);
          A(p, k) = zsub(zscal(cs, apk), zscal(sn, zmul(e, aqk)));
          A(q, k) = zadd(zscal(sn, zmul(zconj(e), apk)), zscal(cs, aqk));
        }
        A(p, q) = (zd){0.0, 0.0};
        A(q, p) = (zd){0.0, 0.0};
        A(p, p).i = 0.0;
        A(q, q).i = 0.0;
      }
  }
  /* ascending order (stable selection) */
  int *idx = (int *)malloc(sizeof(int) * (n ? n : 1));
  for (int i = 0; i < n; i++) idx[i] = i;
  for (int i = 0; i < n; i++) {
    int m = i;
    for (int j = i + 1; j < n; j++)
      if (A(idx[j], idx[j]).r < A(idx[m], idx[m]).r) m = j;
    int t = idx[i]; idx[i] = idx[m]; idx[m] = t;
  }
  for (int j = 0; j < n; j++) w[j] = A(idx[j], idx[j]).r;
  for (int j = 0; j < n; j++) {
    const zd *col = v + (size_t)n * idx[j];
    int im = 0;
    double best = -1.0;
    for (int r = 0; r < n; r++) {
      const double m = hypot(col[r].r, col[r].i);
      if (m > best * (1.0 + 1e-12)) { best = m; im = r; }
    }
    const zd ph = {col[im].r / best, -col[im].i / best}; /* conj(col[im]) / |col[im]| */
    for (int r = 0; r < n; r++) A(r, j) = zmul(col[r], ph);
  }
  free(idx);
  free(v);
  return 0;
#undef A
#undef V
}

/* ------------------------------------------------------------------ rayleigh-ritz8f.F90:128-242, 578-683
 * caqr = dV Psi^H (H Psi) (H Psi from rayleighritz_hamiltonian_c when l_rrb), zheev('v','u'),
 * Psi <- Psi Z, S Psi <- S Psi Z, cspsep <- cspsep Z; hsvcm ends as S Psi Z. */
int orc_rayleighritz_c(int nso, int natom, int num_ppcell, int num_spe, int nprjmx, int neigmx,
                       int numk, int nums, int ncol, int nspv, int num_list, int ncpx, int ncpy,
                       int ncpz, int npxmax, int npymax, int npzmax, int nperi, int nk, int ns1,
                       int nf, int ndisp, int key_natpri_in, int key_natpri_inps,
                       int key_soc_calc, double dx, double dy, double dz, double xmax,
                       double ymax, double zmax, const double *skpxx, const double *skpyy,
                       const double *skpzz, const int *nprj, const int *natinf,
                       const int *lstvec2, const int *indspe, const int *natpri, const int *naps,
                       const int *natsoc, const int *ntyppp, const int *lstx, const int *lsty,
                       const int *lstz, const int *natx, const int *naty, const int *natz,
                       const double *dij, const double *dijsoc, const double *veff,
                       const double *vnlocp, int l_rrb, double *cspsep, double *sveccm,
                       double *ssvcm, double *hsvcm, double *sval) {
  (void)numk; (void)npxmax; (void)npymax; (void)npzmax; (void)ndisp; (void)key_soc_calc;
  (void)xmax; (void)ymax; (void)zmax; (void)natsoc; (void)ntyppp; (void)dijsoc;
  if (ncol != 1 || nso != 0)
    return fail(5, "rayleighritz_c: noncollinear spin and spin-orbit are not part of this build");
  zctx_t c;
  zctx_set(&c, ncpx, ncpy, ncpz, nf, nperi, dx, dy, dz, natom, num_spe, nums, nprjmx, num_list,
           num_ppcell, key_natpri_in, key_natpri_inps, nprj, indspe, natinf, lstvec2, natpri, naps,
           lstx, lsty, lstz, natx, naty, natz, vnlocp, NULL, skpxx[nk - 1], skpyy[nk - 1],
           skpzz[nk - 1]);
  c.dij = dij;
  zctx_alloc(&c);
  const int nsv = ns1 < nspv ? ns1 : nspv;
  c.veff = veff + (size_t)(nsv - 1) * c.n;
  const long n = c.n, n1 = c.n1;
  const int M = neigmx;
  const size_t off = (size_t)n * M * ((ns1 - 1) + (size_t)nums * (nk - 1));
  const size_t off1 = (size_t)n1 * M * ((ns1 - 1) + (size_t)nums * (nk - 1));
  zd *Psi = (zd *)sveccm + off, *SPsi = (zd *)ssvcm + off, *Cps = (zd *)cspsep + off1;
  zd *H = (zd *)hsvcm;
  int rc = 0;
  if (l_rrb)
    for (int l = 1; l <= M && !rc; l++)
      rc = zapply_h(&c, Psi + (size_t)n * (l - 1), Cps + (size_t)n1 * (l - 1), H + (size_t)n * (l - 1));
  zd *A = (zd *)zalloc((size_t)M * M), *T = (zd *)zalloc((size_t)n * M);
  zd *tc = (zd *)zalloc((size_t)n1 * M);
  if (!rc) {
    for (int j = 0; j < M; j++)
      for (int i = 0; i < M; i++)
        A[i + (size_t)M * j] = zscal(c.dv, zdotc(n, Psi + (size_t)n * i, H + (size_t)n * j));
    double *w = sval + (size_t)M * ((ns1 - 1) + (size_t)(nums + 1 - ncol) * (nk - 1));
    rc = orc_zheev(M, (double *)A, w);
    if (rc) rc = fail(3, "rayleighritz_diag_c: error in matrix diagonalization");
  }
  if (!rc) {
    /* Psi <- Psi Z (through hsvcm), S Psi <- S Psi Z (hsvcm keeps it), cspsep <- cspsep Z */
    for (int j = 0; j < M; j++)
      for (long i = 0; i < n; i++) {
        zd s = {0.0, 0.0};
        for (int k = 0; k < M; k++) s = zadd(s, zmul(Psi[i + (size_t)n * k], A[k + (size_t)M * j]));
        T[i + (size_t)n * j] = s;
      }
    memcpy(Psi, T, sizeof(zd) * n * M);
    for (int j = 0; j < M; j++)
      for (long i = 0; i < n; i++) {
        zd s = {0.0, 0.0};
        for (int k = 0; k < M; k++) s = zadd(s, zmul(SPsi[i + (size_t)n * k], A[k + (size_t)M * j]));
        T[i + (size_t)n * j] = s;
      }
    memcpy(SPsi, T, sizeof(zd) * n * M);
    memcpy(H, T, sizeof(zd) * n * M);
    for (int j = 0; j < M; j++)
      for (long i = 0; i < n1; i++) {
        zd s = {0.0, 0.0};
        for (int k = 0; k < M; k++) s = zadd(s, zmul(Cps[i + (size_t)n1 * k], A[k + (size_t)M * j]));
        tc[i + (size_t)n1 * j] = s;
      }
    memcpy(Cps, tc, sizeof(zd) * n1 * M);
  }
  free(A); free(T); free(tc);
  zctx_free(&c);
  return rc;
}

/* ------------------------------------------------------------------ reorderstates8f.F90:70-124 (nrc = 1) */
int orc_reorderstates_c(int num_ppcell, int nprjmx, int neigmx, int nums, int ncol, int ncpx,
                        int ncpy, int ncpz, int ns1, int nk, double *sval, double *residual_states,
                        double *cspsep, double *sveccm, double *ssvcm, double *hsvcm) {
  const int M = neigmx, d2 = nums + 1 - ncol;
  const long n = (long)ncpx * ncpy * ncpz, n1 = (long)nprjmx * num_ppcell;
  double *sv = sval + (size_t)M * ((ns1 - 1) + (size_t)d2 * (nk - 1));
  double *rs = residual_states + (size_t)M * ((ns1 - 1) + (size_t)d2 * (nk - 1));
  const size_t off = (size_t)n * M * ((ns1 - 1) + (size_t)nums * (nk - 1));
  const size_t off1 = (size_t)n1 * M * ((ns1 - 1) + (size_t)nums * (nk - 1));
  zd *P = (zd *)sveccm + off, *SP = (zd *)ssvcm + off, *C = (zd *)cspsep + off1, *Hh = (zd *)hsvcm;
  int *lank = (int *)calloc(M ? M : 1, sizeof(int));
  double *st = (double *)calloc(M ? M : 1, sizeof(double)), *rt = (double *)calloc(M ? M : 1, sizeof(double));
  zd *cp = (zd *)zalloc((size_t)n1 * M);
  lank[0] = 1;
  for (int j = 2; j <= M; j++) {
    lank[j - 1] = 1;
    for (int i = 1; i <= j - 1; i++) {
      if (sv[i - 1] > sv[j - 1]) lank[i - 1]++;
      else lank[j - 1]++;
    }
  }
  for (int j = 1; j <= M; j++) {
    st[lank[j - 1] - 1] = sv[j - 1];
    rt[lank[j - 1] - 1] = rs[j - 1];
    memcpy(cp + (size_t)n1 * (lank[j - 1] - 1), C + (size_t)n1 * (j - 1), sizeof(zd) * n1);
    memcpy(Hh + (size_t)n * (lank[j - 1] - 1), P + (size_t)n * (j - 1), sizeof(zd) * n);
  }
  for (int j = 1; j <= M; j++) {
    sv[j - 1] = st[j - 1];
    rs[j - 1] = rt[j - 1];
  }
  memcpy(C, cp, sizeof(zd) * n1 * M);
  memcpy(P, Hh, sizeof(zd) * n * M);
  for (int j = 1; j <= M; j++)
    memcpy(Hh + (size_t)n * (lank[j - 1] - 1), SP + (size_t)n * (j - 1), sizeof(zd) * n);
  memcpy(SP, Hh, sizeof(zd) * n * M);
  free(lank); free(st); free(rt); free(cp);
  return 0;
}

/* ------------------------------------------------------------------ scf_charge8f.F90:324-381 (ncol = 1)
 * rhosmt(:,ns) = sum_{k,l} fnele(l,ns,k) |psi|^2 */
int orc_scf_charge_smt_c(int neigmx, int nums, int ncol, int nspv, int numk, int ncpx, int ncpy,
                         int ncpz, const double *fnele, const double *sveccm, double *rhosmt) {
  if (ncol != 1) return fail(5, "scf_charge_smt_c: noncollinear spin (ncol = 2) is not part of this build");
  const long n = (long)ncpx * ncpy * ncpz;
  for (long i = 0; i < n * nspv; i++) rhosmt[i] = 0.0;
  for (int nk = 1; nk <= numk; nk++)
    for (int ns = 1; ns <= nums; ns++) {
      double *rho = rhosmt + (size_t)n * (nspv > 1 ? ns - 1 : 0);
      for (int l = 1; l <= neigmx; l++) {
        const double f = fnele[(l - 1) + (size_t)neigmx * ((ns - 1) + (size_t)nums * (nk - 1))];
        const zd *p = (const zd *)sveccm + (size_t)n * ((l - 1) + (size_t)neigmx * ((ns - 1) + (size_t)nums * (nk - 1)));
        for (long i = 0; i < n; i++) rho[i] = rho[i] + f * (p[i].r * p[i].r + p[i].i * p[i].i);
      }
    }
  return 0;
}

/* ------------------------------------------------------------------ ksdiis8f.F90:715-1509
 * ksdiis_c: RMM-DIIS of H_k for spin ns1, k-point nk (ncol = 1, nso = 0).  The reference
 * works through the bands in blocks of nb0 = 64 for its MPI reductions; every operation inside
 * a block acts on each band on its own, so the bands are taken one at a time here.  With
 * nitermax = 2 (:738) the inner loop runs once (iiter = 1).  The DIIS combination's overall
 * phase is the Hermitian eigensolver's free choice in the reference (zheev, :1216, :1262);
 * here the coefficient of the incoming band is made real and non-negative. */
int orc_ksdiis_c(int l_hsv, int nso, int natom, int num_spe, int neigmx, int nums, int ncol,
                 int nspv, int numk, int nperi, int ndiismax, int nprecon_diis, int nk, int ns1,
                 int nf, int nprjmx, int num_list, int num_ppcell, int ndisp, int northo, int ncpx,
                 int ncpy, int ncpz, int npxmax, int npymax, int npzmax,
                 int key_ortho_cmpt_innerproduct, int key_natpri_in, int key_natpri_inps,
                 int key_soc_calc, double eps_eig_diis, double alambda_diis, double ratio_diis,
                 double alambda_min, double alambda_max, double epssd, double xmax, double ymax,
                 double zmax, double skpx, double skpy, double skpz, const int *indspe,
                 const int *natpri, const int *naps, const int *nprj, const int *natinf,
                 const int *lstvec2, const int *latom, const int *natsoc, const int *ntyppp,
                 const int *lstx, const int *lsty, const int *lstz, const int *natx, const int *naty,
                 const int *natz, const double *dij, const double *dijsoc, const double *veff,
                 const double *vnlocp, const double *sss, int *ksconv, int *ksitmax, double *cspsep,
                 double *sval, double *sveccm, double *ssvcm, double *residual_states,
                 double *hsvcm) {
  (void)numk; (void)ndisp; (void)npxmax; (void)npymax; (void)npzmax; (void)key_soc_calc;
  (void)ratio_diis; (void)latom; (void)natsoc; (void)ntyppp; (void)dijsoc;
  if (ncol != 1 || nso != 0)
    return fail(5, "ksdiis_c: noncollinear spin and spin-orbit are not part of this build");
  const double dx = 2.0 * xmax / ncpx, dy = 2.0 * ymax / ncpy, dz = 2.0 * zmax / ncpz;
  zctx_t c;
  zctx_set(&c, ncpx, ncpy, ncpz, nf, nperi, dx, dy, dz, natom, num_spe, nums, nprjmx, num_list,
           num_ppcell, key_natpri_in, key_natpri_inps, nprj, indspe, natinf, lstvec2, natpri, naps,
           lstx, lsty, lstz, natx, naty, natz, vnlocp, sss, skpx, skpy, skpz);
  c.xmax = xmax; c.ymax = ymax; c.zmax = zmax;
  c.dij = dij;
  zctx_alloc(&c);
  const int nsv = ns1 < nspv ? ns1 : nspv;
  c.veff = veff + (size_t)(nsv - 1) * c.n;
  const long n = c.n, n1 = c.n1;
  const size_t off = (size_t)n * neigmx * ((ns1 - 1) + (size_t)nums * (nk - 1));
  const size_t off1 = (size_t)n1 * neigmx * ((ns1 - 1) + (size_t)nums * (nk - 1));
  zd *Psi = (zd *)sveccm + off, *SPsi = (zd *)ssvcm + off, *Cps = (zd *)cspsep + off1;
  zd *Hps = (zd *)hsvcm;
  double *sv = sval + (size_t)neigmx * ((ns1 - 1) + (size_t)nums * (nk - 1));
  enum { NIT = 2 };
  zd *phi = (zd *)zalloc(n * (NIT + 1)), *sphi = (zd *)zalloc(n * (NIT + 1));
  zd *hphi = (zd *)zalloc(n * (NIT + 1)), *rvec = (zd *)zalloc(n * (NIT + 1));
  zd *cphi = (zd *)zalloc(n1 * (NIT + 1));
  zd *rv = (zd *)zalloc(n), *av = (zd *)zalloc(n), *asv = (zd *)zalloc(n), *cp = (zd *)zalloc(n1);
#define PH(j) (phi + (size_t)n * (j))
#define SPH(j) (sphi + (size_t)n * (j))
#define HPH(j) (hphi + (size_t)n * (j))
#define RV(j) (rvec + (size_t)n * (j))
#define CPH(j) (cphi + (size_t)n1 * (j))
  int rc = 0, jiter = 0;
  while (jiter + 1 < ndiismax && !rc) {
    jiter++;
    for (int l = 1; l <= neigmx && !rc; l++) {
      memcpy(PH(0), Psi + (size_t)n * (l - 1), sizeof(zd) * n);
      memcpy(SPH(0), SPsi + (size_t)n * (l - 1), sizeof(zd) * n);
      memcpy(CPH(0), Cps + (size_t)n1 * (l - 1), sizeof(zd) * n1);
      const int iiter = 1;
      /* eigenvalue and residual (:868-917): H phi_0 with the stored <p|phi_0> */
      memcpy(cp, CPH(0), sizeof(zd) * n1);
      if ((rc = zapply_h(&c, PH(0), cp, HPH(0)))) break;
      const double valre = zdotr(n, PH(0), HPH(0)) * c.dv;
      sv[l - 1] = valre;
      for (long i = 0; i < n; i++) RV(0)[i] = zadd(HPH(0)[i], zscal(-valre, SPH(0)[i]));
      /* precondition (:920-954), <p|K r>, H K r, S K r (:956-1063) */
      if (nprecon_diis == 1) zprecond(&c, RV(0), rv);
      else memcpy(rv, RV(0), sizeof(zd) * n);
      zproj(&c, rv, cp);
      if ((rc = zapply_h(&c, rv, cp, av))) break;
      zapply_s(&c, rv, cp, asv);
      /* step length from a trial step alambda_diis (:1065-1096) */
      const double a1 = alambda_diis;
      for (long i = 0; i < n1; i++) CPH(1)[i] = zadd(CPH(0)[i], zscal(a1, cp[i]));
      for (long i = 0; i < n; i++) PH(1)[i] = zadd(PH(0)[i], zscal(a1, rv[i]));
      for (long i = 0; i < n; i++) HPH(1)[i] = zadd(HPH(0)[i], zscal(a1, av[i]));
      for (long i = 0; i < n; i++) SPH(1)[i] = zadd(SPH(0)[i], zscal(a1, asv[i]));
      const double s1 = zdotr(n, rv, HPH(0)), s2 = zdotr(n, rv, SPH(0));
      const double s3 = zdotr(n, PH(1), HPH(1)), s4 = zdotr(n, PH(1), SPH(1));
      const double aaa = 2.0 * (s1 - sv[l - 1] * s2) * c.dv;
      const double bbb = (s3 / s4 - sv[l - 1] - aaa * a1) / (a1 * a1);
      double alambda = 0.5 * aaa / bbb;
      if (alambda > alambda_max) alambda = alambda_max;
      if (alambda < alambda_min) alambda = alambda_min;
      /* trial phi (:1099-1116) and residuals (:1140-1149) */
      for (long i = 0; i < n; i++) PH(iiter)[i] = zadd(PH(0)[i], zscal(alambda, rv[i]));
      for (long i = 0; i < n; i++) HPH(iiter)[i] = zadd(HPH(0)[i], zscal(alambda, av[i]));
      for (long i = 0; i < n; i++) SPH(iiter)[i] = zadd(SPH(0)[i], zscal(alambda, asv[i]));
      for (long i = 0; i < n1; i++) CPH(iiter)[i] = zadd(CPH(0)[i], zscal(alambda, cp[i]));
      for (int j = 0; j <= iiter; j++)
        for (long i = 0; i < n; i++) RV(j)[i] = zadd(HPH(j)[i], zscal(-sv[l - 1], SPH(j)[i]));
      /* DIIS matrices (:1154-1199): csmat(i,j) = dV <phi_i|S phi_j>, crmat(i,j) = dV <r_i|r_j> */
      const int m = iiter + 1;
      zd csm[(NIT + 1) * (NIT + 1)], crm[(NIT + 1) * (NIT + 1)], csm0[(NIT + 1) * (NIT + 1)];
      zd cvect[(NIT + 1) * (NIT + 1)], csm2[(NIT + 1) * (NIT + 1)];
      double eigen[NIT + 1], work[NIT + 1];
#define M2(a, i, j) a[((i)-1) + m * ((j)-1)]
      for (int j = 0; j <= iiter; j++)
        for (int i = 0; i <= j; i++) {
          const zd s = zscal(c.dv, zdotc(n, PH(i), SPH(j))), r = zscal(c.dv, zdotc(n, RV(i), RV(j)));
          M2(csm, i + 1, j + 1) = s; M2(csm, j + 1, i + 1) = zconj(s);
          M2(crm, i + 1, j + 1) = r; M2(crm, j + 1, i + 1) = zconj(r);
        }
      const double tmp = 1.0 / sqrt(M2(csm, m, m).r);
      for (long i = 0; i < n; i++) {
        PH(iiter)[i] = zscal(tmp, PH(iiter)[i]); HPH(iiter)[i] = zscal(tmp, HPH(iiter)[i]);
        SPH(iiter)[i] = zscal(tmp, SPH(iiter)[i]); RV(iiter)[i] = zscal(tmp, RV(iiter)[i]);
      }
      for (long i = 0; i < n1; i++) CPH(iiter)[i] = zscal(tmp, CPH(iiter)[i]);
      for (int i = 1; i <= m; i++) { M2(csm, m, i) = zscal(tmp, M2(csm, m, i)); M2(crm, m, i) = zscal(tmp, M2(crm, m, i)); }
      for (int i = 1; i <= m; i++) { M2(csm, i, m) = zscal(tmp, M2(csm, i, m)); M2(crm, i, m) = zscal(tmp, M2(crm, i, m)); }
      memcpy(csm0, csm, sizeof csm);
      /* overlap eigenvectors, descending (:1216-1225) */
      memcpy(csm2, csm, sizeof csm);
      if ((rc = orc_zheev(m, (double *)csm2, eigen))) break;
      for (int i = 1; i <= m; i++) work[i - 1] = eigen[i - 1];
      for (int j = 1; j <= m; j++) {
        for (int i = 1; i <= m; i++) M2(cvect, i, j) = M2(csm2, i, m + 1 - j);
        eigen[j - 1] = work[m + 1 - j - 1];
      }
      /* drop overlap eigenvalues below eps_eig_diis (:1227-1240) */
      int nmat = 0, k = 0;
      while (k < m) {
        nmat++;
        k++;
        if (eigen[nmat - 1] < eps_eig_diis) {
          for (int i = nmat; i <= iiter - 1; i++) {
            for (int r = 1; r <= m; r++) M2(cvect, r, i) = M2(cvect, r, i + 1);
            eigen[i - 1] = eigen[i];
          }
          nmat--;
        }
      }
      zd coef[NIT + 1];
      if (nmat > 1) {
        for (int j = 1; j <= nmat; j++)
          for (int i = 1; i <= m; i++) M2(cvect, i, j) = zscal(1.0 / sqrt(eigen[j - 1]), M2(cvect, i, j));
        /* projected residual matrix and its lowest eigenvector (:1242-1275) */
        zd t1[(NIT + 1) * (NIT + 1)], cr2[(NIT + 1) * (NIT + 1)], cv2[(NIT + 1) * (NIT + 1)];
        for (int kk = 1; kk <= m; kk++)
          for (int j = 1; j <= nmat; j++) {
            zd s = {0.0, 0.0};
            for (int i = 1; i <= m; i++) s = zadd(s, zmul(M2(crm, kk, i), M2(cvect, i, j)));
            M2(t1, kk, j) = s;
          }
        for (int kk = 1; kk <= nmat; kk++)
          for (int j = 1; j <= nmat; j++) {
            zd s = {0.0, 0.0};
            for (int i = 1; i <= m; i++) s = zadd(s, zmul(zconj(M2(cvect, i, kk)), M2(t1, i, j)));
            cr2[(kk - 1) + nmat * (j - 1)] = s;
          }
        if ((rc = orc_zheev(nmat, (double *)cr2, eigen))) break;
        for (int i = 1; i <= nmat; i++) work[i - 1] = eigen[i - 1];
        for (int j = 1; j <= nmat; j++) {
          for (int i = 1; i <= nmat; i++) cv2[(i - 1) + nmat * (j - 1)] = cr2[(i - 1) + nmat * (nmat - j)];
          eigen[j - 1] = work[nmat - j];
        }
        residual_states[l - 1] = eigen[nmat - 1];
        /* the combination and its S-normalisation (:1293-1310) */
        for (int j = 1; j <= m; j++) {
          zd s = {0.0, 0.0};
          for (int i = 1; i <= nmat; i++) s = zadd(s, zmul(M2(cvect, j, i), cv2[(i - 1) + nmat * (nmat - 1)]));
          coef[j - 1] = s;
        }
        double t2 = 0.0;
        for (int j = 1; j <= m; j++)
          for (int i = 1; i <= m; i++) t2 += zmul(zmul(zconj(coef[i - 1]), M2(csm0, i, j)), coef[j - 1]).r;
        for (int j = 1; j <= m; j++) coef[j - 1] = zscal(1.0 / sqrt(t2), coef[j - 1]);
      } else {
        for (int j = 1; j <= m; j++) coef[j - 1] = (zd){0.0, 0.0};
        coef[m - 1] = (zd){1.0, 0.0};
      }
#undef M2
      /* phase: the incoming band's coefficient real and non-negative */
      const double a0 = hypot(coef[0].r, coef[0].i);
      if (a0 > 0.0) {
        const zd ph = {coef[0].r / a0, -coef[0].i / a0};
        for (int j = 0; j < m; j++) coef[j] = zmul(coef[j], ph);
        coef[0].i = 0.0;
      }
      /* new band (:1313-1352; ksdiis_c_03 adds the terms in pairs) */
      zd *outs[4] = {Psi + (size_t)n * (l - 1), SPsi + (size_t)n * (l - 1), Hps + (size_t)n * (l - 1), NULL};
      const zd *srcs[3] = {phi, sphi, hphi};
      for (int q = 0; q < 3; q++)
        for (long i = 0; i < n; i++) {
          zd s = zadd(zmul(coef[0], srcs[q][i]), zmul(coef[1], srcs[q][i + n]));
          outs[q][i] = s;
        }
      for (long i = 0; i < n1; i++) {
        zd s = {0.0, 0.0};
        for (int j = 0; j < m; j++) s = zadd(s, zmul(coef[j], CPH(j)[i]));
        Cps[(size_t)n1 * (l - 1) + i] = s;
      }
      if (iiter > *ksitmax) *ksitmax = iiter;
      *ksconv = *ksconv && (sqrt(residual_states[l - 1]) <= epssd);
    }
    if (rc) break;
    /* l_hsv: orthonormalise, then H psi of every band into hsvcm (:1416-1478) */
    if (l_hsv) {
      for (int i = 1; i <= northo && !rc; i++)
        rc = zortho_c_01(&c, neigmx, 0, key_ortho_cmpt_innerproduct, key_ortho_cmpt_innerproduct,
                         Psi, SPsi, Cps);
      for (int l = 1; l <= neigmx && !rc; l++) {
        memcpy(cp, Cps + (size_t)n1 * (l - 1), sizeof(zd) * n1);
        rc = zapply_h(&c, Psi + (size_t)n * (l - 1), cp, Hps + (size_t)n * (l - 1));
      }
    }
  }
#undef PH
#undef SPH
#undef HPH
#undef RV
#undef CPH
  free(phi); free(sphi); free(hphi); free(rvec); free(cphi);
  free(rv); free(av); free(asv); free(cp);
  zctx_free(&c);
  return rc;
}
