/*
 * kukan_oracle_ckpt.c — CPU restatement of the RSPACE kukan v8f checkpoint files (rw_wfpot8f.F90):
 * wfc.* (wave functions, version 5), rho.* / r_o.* (densities, version 3) and vht.* (dense
 * Hartree potential, version 3), written and read as Fortran sequential unformatted files (each
 * record framed by its 4-byte little-endian byte count), one rank.
 *
 * TEST INFRASTRUCTURE ONLY (the checker), like kukan_oracle.c: loaded by the tests, never by the
 * product library.  Every routine follows the reference routine cited above it (file:line relative
 * to RSPACEes.v08.01/kukan/).
 *
 * Pins (tests/test_reference_ckpt.py): files written here are byte-identical to the reference's
 * (write_wf, write_rhovht with every nevhist), and the readers reproduce the reference's arrays
 * from reference-written files in every mapping case of read_wf (same shape; fewer spins in the
 * file; real file into a complex run; a collinear two-spin file into a noncollinear run; half the
 * k-points with inversion symmetry) and of read_rho (spin expansion), and its error cases.
 */
#include "kukan_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int orc_set_error(int code, const char *msg); /* kukan_oracle.c */

enum { WFC_VERSION = 5, RHO_VERSION = 3, VHT_VERSION = 3 };

/* ------------------------------------------------------------------ rw_wfpot8f.F90:535-572 get_fname
 * "<chnam>.<nr1:5 digits>[.<nr2:3 digits>]", prefixed by "/" when chdir is not blank; the file
 * is chdir // fname. */
int orc_ckpt_path(const char *chdir, const char *chnam, int nr1, int nr2, char *out, int len) {
  if (nr1 >= 100000 || (nr2 >= 0 && nr2 >= 1000)) return orc_set_error(1, "get_fname: not enough digits");
  char cd[1024];
  snprintf(cd, sizeof cd, "%s", chdir ? chdir : "");
  size_t k = strlen(cd);
  while (k > 0 && cd[k - 1] == ' ') cd[--k] = 0; /* trim */
  int w;
  if (nr2 >= 0) w = snprintf(out, (size_t)len, "%s%s%s.%05d.%03d", cd, k ? "/" : "", chnam, nr1, nr2);
  else w = snprintf(out, (size_t)len, "%s%s%s.%05d", cd, k ? "/" : "", chnam, nr1);
  if (w < 0 || w >= len) return orc_set_error(1, "get_fname: string too short");
  return 0;
}

/* sequential unformatted records */
static int put_rec(FILE *f, const void *const *parts, const size_t *bytes, int nparts) {
  size_t tot = 0;
  for (int i = 0; i < nparts; i++) tot += bytes[i];
  if (tot > 0x7fffffffUL) return orc_set_error(1, "record longer than 2 GiB");
  const int m = (int)tot;
  if (fwrite(&m, 4, 1, f) != 1) return orc_set_error(1, "write error");
  for (int i = 0; i < nparts; i++)
    if (bytes[i] && fwrite(parts[i], 1, bytes[i], f) != bytes[i]) return orc_set_error(1, "write error");
  if (fwrite(&m, 4, 1, f) != 1) return orc_set_error(1, "write error");
  return 0;
}
static int put1(FILE *f, const void *p, size_t b) {
  const void *parts[1] = {p};
  return put_rec(f, parts, &b, 1);
}
/* read the first `want` bytes of the next record (fewer bytes in the record: error) */
static int get_rec(FILE *f, void *dst, size_t want, const char *err) {
  int m = 0, m2 = 0;
  if (fread(&m, 4, 1, f) != 1 || m < 0 || (size_t)m < want) return orc_set_error(1, err);
  if (want && fread(dst, 1, want, f) != want) return orc_set_error(1, err);
  if (fseek(f, (long)((size_t)m - want), SEEK_CUR) != 0) return orc_set_error(1, err);
  if (fread(&m2, 4, 1, f) != 1 || m2 != m) return orc_set_error(1, err);
  return 0;
}

/* ------------------------------------------------------------------ rw_wfpot8f.F90:352-399 write_wf */
int orc_write_wf(int nr1In, int nr2In, const char *chdir, int nrc, int nums, int ncol, int neigmx,
                 int numk, int ncpx, int ncpy, int ncpz, int nrxIn, int nryIn, int nrzIn, int ksym,
                 const double *svecre, const double *sveccm, const double *sval) {
  const int nr1 = nr1In < 0 ? 0 : nr1In, nr2 = nr2In < 0 ? 0 : nr2In;
  const int nrx = nrxIn < 0 ? 0 : nrxIn, nry = nryIn < 0 ? 0 : nryIn, nrz = nrzIn < 0 ? 0 : nrzIn;
  char path[1100];
  int rc = orc_ckpt_path(chdir, "wfc", nr1, nr2, path, sizeof path);
  if (rc) return rc;
  FILE *f = fopen(path, "wb");
  if (!f) return orc_set_error(1, "cannot open wfc.* file(s)");
  const int ver = WFC_VERSION, dims[8] = {nrc, nums, ncol, neigmx, numk, ncpx, ncpy, ncpz};
  const int ranks[4] = {nrx, nry, nrz, 0};
  const size_t n = (size_t)ncpx * ncpy * ncpz, el = nrc ? 16 : 8;
  rc = put1(f, &ver, 4);
  if (!rc) rc = put1(f, dims, sizeof dims);
  if (!rc) rc = put1(f, ranks, sizeof ranks);
  if (!rc) rc = put1(f, &ksym, 4);
  if (!rc) rc = put1(f, sval, 8 * (size_t)neigmx * (nums + 1 - ncol) * numk);
  const char *v = (const char *)(nrc ? sveccm : svecre);
  for (size_t b = 0; b < (size_t)neigmx * nums * numk && !rc; b++) rc = put1(f, v + el * n * b, el * n);
  fclose(f);
  return rc;
}

/* ------------------------------------------------------------------ rw_wfpot8f.F90:7-185 read_wf */
int orc_read_wf(int nr1In, int nr2In, const char *chdir, int nrc, int nums, int ncol, int neigmx,
                int numk, int ncpx, int ncpy, int ncpz, int ksym, int key_ksym_inv2, double *svecre,
                double *sveccm, double *sval) {
  const int nr1 = nr1In < 0 ? 0 : nr1In, nr2 = nr2In < 0 ? 0 : nr2In;
  const char *E = "error found in wfc.*";
  char path[1100];
  int rc = orc_ckpt_path(chdir, "wfc", nr1, nr2, path, sizeof path);
  if (rc) return rc;
  FILE *f = fopen(path, "rb");
  if (!f) return orc_set_error(1, "cannot open wfc.* file(s)");
  int ver = 0, d[8], ksym1 = 0;
  if ((rc = get_rec(f, &ver, 4, E))) goto out;
  if (ver != WFC_VERSION) { rc = orc_set_error(1, "wrong wfc version"); goto out; }
  if ((rc = get_rec(f, d, sizeof d, E)) || (rc = get_rec(f, NULL, 0, E)) || (rc = get_rec(f, &ksym1, 4, E))) goto out;
  {
    const int nrc1 = d[0], nums1 = d[1], ncol1 = d[2], neigmx1 = d[3], numk1 = d[4];
    const int ncpx1 = d[5], ncpy1 = d[6], ncpz1 = d[7];
    static char msg[1400];
    if (ncpx1 != ncpx || ncpy1 != ncpy || ncpz1 != ncpz || nrc1 > nrc || nums1 > nums || ncol1 > ncol ||
        (numk1 != numk && (numk1 * 2 != numk || ksym != key_ksym_inv2))) {
      snprintf(msg, sizeof msg, "read_wf: unexpected array dimensions in file %s", path);
      rc = orc_set_error(1, msg);
      goto out;
    }
    if (neigmx1 * ncol < neigmx * ncol1) {
      snprintf(msg, sizeof msg, "read_wf: not enough eigenstates in file %s", path);
      rc = orc_set_error(1, msg);
      goto out;
    }
    const size_t n = (size_t)ncpx * ncpy * ncpz;
    const int d2 = nums + 1 - ncol;
#define SV(l, ns, nk) sval[((l) - 1) + (size_t)neigmx * (((ns) - 1) + (size_t)d2 * ((nk) - 1))]
#define VR(l, ns, nk) (svecre + n * (((l) - 1) + (size_t)neigmx * (((ns) - 1) + (size_t)nums * ((nk) - 1))))
#define VC(l, ns, nk) (sveccm + 2 * n * (((l) - 1) + (size_t)neigmx * (((ns) - 1) + (size_t)nums * ((nk) - 1))))
    if (nrc1 == nrc && nums1 == nums && ncol1 == ncol && neigmx1 == neigmx && numk1 == numk) {
      if ((rc = get_rec(f, sval, 8 * (size_t)neigmx * d2 * numk, E))) goto out;
      for (int nk = 1; nk <= numk; nk++)
        for (int ns = 1; ns <= nums; ns++)
          for (int l = 1; l <= neigmx; l++)
            if ((rc = get_rec(f, nrc ? VC(l, ns, nk) : VR(l, ns, nk), (nrc ? 16 : 8) * n, E))) goto out;
      goto out;
    }
    const int d21 = nums1 + 1 - ncol1;
    double *sval1 = (double *)calloc((size_t)neigmx1 * d21 * numk1, 8);
    double *vr1 = (double *)calloc(n, 8), *vc1 = (double *)calloc(2 * n, 8);
#define SV1(l, ns, nk) sval1[((l) - 1) + (size_t)neigmx1 * (((ns) - 1) + (size_t)d21 * ((nk) - 1))]
    rc = get_rec(f, sval1, 8 * (size_t)neigmx1 * d21 * numk1, E);
    if (!rc && (ncol1 == ncol || nums1 < nums)) {
      for (int nk = 1; nk <= numk1 && !rc; nk++)
        for (int ns = 1; ns <= nums1 && !rc; ns++)
          for (int l1 = 1; l1 <= neigmx1 && !rc; l1++) {
            const int l = ncol1 == ncol ? l1 : 2 * l1 - 1;
            if (l <= neigmx && ns <= d2) SV(l, ns, nk) = SV1(l1, ns, nk);
            if ((rc = get_rec(f, nrc1 ? (void *)vc1 : (void *)vr1, (nrc1 ? 16 : 8) * n, E))) break;
            if (l > neigmx) continue;
            if (nrc1 < nrc)
              for (size_t i = 0; i < n; i++) { vc1[2 * i] = vr1[i]; vc1[2 * i + 1] = 0.0; }
            if (nrc == 0) memcpy(VR(l, ns, nk), vr1, 8 * n);
            else memcpy(VC(l, ns, nk), vc1, 16 * n);
            if (nums1 < nums) {
              if (ncol1 == ncol) {
                SV(l, nums, nk) = SV(l, 1, nk);
                if (nrc == 0) memcpy(VR(l, nums, nk), VR(l, 1, nk), 8 * n);
                else memcpy(VC(l, nums, nk), VC(l, 1, nk), 16 * n);
              } else {
                memset(VC(l, nums, nk), 0, 16 * n);
                if (l + 1 <= neigmx) {
                  SV(l + 1, 1, nk) = SV(l, 1, nk);
                  memcpy(VC(l + 1, nums, nk), VC(l, 1, nk), 16 * n);
                  memset(VC(l + 1, 1, nk), 0, 16 * n);
                }
              }
            }
          }
    } else if (!rc) {
      /* a collinear two-spin file into a noncollinear run: bands merged by energy */
      int *map = (int *)calloc((size_t)nums1 * neigmx1, sizeof(int));
#define MAP(ns, l1) map[((ns) - 1) + (size_t)nums1 * ((l1) - 1)]
      for (int nk = 1; nk <= numk1 && !rc; nk++) {
        memset(map, 0, sizeof(int) * (size_t)nums1 * neigmx1);
        int lspin[2] = {1, 1};
        for (int l = 1; l <= neigmx; l++) {
          int ns = 0;
          if (lspin[0] > neigmx1) ns = 2;
          if (lspin[1] > neigmx1) ns = 1;
          if (ns == 0) ns = SV1(lspin[0], 1, nk) <= SV1(lspin[1], 2, nk) ? 1 : 2;
          MAP(ns, lspin[ns - 1]) = l;
          lspin[ns - 1]++;
        }
        for (int ns = 1; ns <= nums1 && !rc; ns++)
          for (int l1 = 1; l1 <= neigmx1 && !rc; l1++) {
            const int l = MAP(ns, l1);
            if (nrc1 == 0) {
              if ((rc = get_rec(f, vr1, 8 * n, E))) break;
              for (size_t i = 0; i < n; i++) { vc1[2 * i] = vr1[i]; vc1[2 * i + 1] = 0.0; }
            } else if ((rc = get_rec(f, vc1, 16 * n, E))) {
              break;
            }
            if (l > 0) {
              SV(l, 1, nk) = SV1(l1, ns, nk);
              memcpy(VC(l, ns, nk), vc1, 16 * n);
              memset(VC(l, 3 - ns, nk), 0, 16 * n);
            }
          }
      }
#undef MAP
      free(map);
    }
    if (!rc && numk1 * 2 == numk) {
      if (ksym != key_ksym_inv2) rc = orc_set_error(1, "read_wf: wrong ksym for k-point mapping");
      for (int nk = 1; nk <= numk1 && !rc; nk++) {
        const int nk2 = 2 * numk1 + 1 - nk;
        for (int ns = 1; ns <= d2; ns++)
          for (int l = 1; l <= neigmx; l++) SV(l, ns, nk2) = SV(l, ns, nk);
        for (int ns = 1; ns <= nums; ns++)
          for (int l = 1; l <= neigmx; l++) {
            if (nrc == 0) {
              memcpy(VR(l, ns, nk2), VR(l, ns, nk), 8 * n);
            } else {
              double *dst = VC(l, ns, nk2);
              const double *src = VC(l, ns, nk);
              for (size_t i = 0; i < n; i++) { dst[2 * i] = src[2 * i]; dst[2 * i + 1] = -src[2 * i + 1]; }
            }
          }
      }
    }
#undef SV1
    free(sval1); free(vr1); free(vc1);
  }
#undef SV
#undef VR
#undef VC
out:
  fclose(f);
  return rc;
}

/* ------------------------------------------------------------------ rw_wfpot8f.F90:188-302 read_rho */
int orc_read_rho(int nrIn, const char *chdir, int nevhist, int nspv, int ncpx, int ncpy, int ncpz,
                 int nradmx, int npoint, int num_atcell, int *lbrydn, int *na_atcell,
                 int *natcell_inf, double *rhosmt, double *rhotrur, double *rhosmtr) {
  int histyn = nevhist == 1 || nevhist == 3 || nevhist == -1 || nevhist == 4;
  const char *E = "error found in rho.*";
  int rc = 0;
  if (histyn) {
    char path[1100];
    if ((rc = orc_ckpt_path(chdir, nevhist == 4 ? "r_o" : "rho", nrIn < 0 ? 0 : nrIn, -1, path, sizeof path))) return rc;
    FILE *f = fopen(path, "rb");
    histyn = f != NULL;
    if (nevhist == -1 && !histyn) return orc_set_error(1, "read_rho: nevhist=-1, but no density file rho.*");
    if (histyn) {
      *lbrydn = 1;
      int ver = 0, d[7];
      const size_t n = (size_t)ncpx * ncpy * ncpz, ra = (size_t)nradmx * npoint;
      if ((rc = get_rec(f, &ver, 4, E))) goto done;
      if (ver != RHO_VERSION) { rc = orc_set_error(1, "wrong rho version"); goto done; }
      if ((rc = get_rec(f, d, sizeof d, E)) || (rc = get_rec(f, NULL, 0, E))) goto done;
      {
        const int nspv1 = d[0], num_atcell1 = d[6];
        int *rec = (int *)calloc((size_t)num_atcell1 + 1, sizeof(int));
        if ((rc = get_rec(f, rec, 4 * ((size_t)num_atcell1 + 1), E))) { free(rec); goto done; }
        *na_atcell = rec[0];
        static char msg[1400];
        if (nspv1 > nspv && nspv1 == 2) rc = orc_set_error(1, "read_rho:  mapping of (mag dens) to (non-mag dens) not implemented");
        else if (nspv1 > nspv && nspv1 == 4) rc = orc_set_error(1, "read_rho:  mapping of (noco dens) to (col dens) not implemented");
        else if (nspv1 > nspv || d[1] != ncpx || d[2] != ncpy || d[3] != ncpz || d[4] != nradmx || d[5] != npoint) {
          snprintf(msg, sizeof msg, "read_rho: unexpected array dimensions in density file %s", path);
          rc = orc_set_error(1, msg);
        } else if (*na_atcell > num_atcell) {
          rc = orc_set_error(1, "read_rho:  num_atcell is too small");
        }
        if (!rc)
          for (int i = 0; i < *na_atcell; i++) natcell_inf[i] = rec[1 + i];
        free(rec);
        if (rc) goto done;
        if (nspv1 == nspv) {
          if ((rc = get_rec(f, rhosmt, 8 * n * nspv, E))) goto done;
          if (num_atcell1 <= num_atcell) {
            if ((rc = get_rec(f, rhotrur, 8 * ra * nspv * num_atcell1, E))) goto done;
            rc = get_rec(f, rhosmtr, 8 * ra * nspv * num_atcell1, E);
          } else {
            double *t = (double *)calloc(ra * nspv * num_atcell1, 8);
            rc = get_rec(f, t, 8 * ra * nspv * num_atcell1, E);
            if (!rc) memcpy(rhotrur, t, 8 * ra * nspv * (size_t)*na_atcell);
            if (!rc) rc = get_rec(f, t, 8 * ra * nspv * num_atcell1, E);
            if (!rc) memcpy(rhosmtr, t, 8 * ra * nspv * (size_t)*na_atcell);
            free(t);
          }
        } else {
          double *s1 = (double *)calloc(n * nspv1, 8), *t1 = (double *)calloc(ra * nspv1 * num_atcell1, 8);
          double *u1 = (double *)calloc(ra * nspv1 * num_atcell1, 8);
          rc = get_rec(f, s1, 8 * n * nspv1, E);
          if (!rc) rc = get_rec(f, t1, 8 * ra * nspv1 * num_atcell1, E);
          if (!rc) rc = get_rec(f, u1, 8 * ra * nspv1 * num_atcell1, E);
          for (int ns = 1; ns <= nspv && !rc; ns++) {
            const int ns1 = ns < nspv1 ? ns : nspv1;
            double fctr = 1.0;
            if (nspv1 == 1) fctr = 0.5;
            if (ns > 2) fctr = 0.0;
            for (size_t i = 0; i < n; i++) rhosmt[i + n * (ns - 1)] = s1[i + n * (ns1 - 1)] * fctr;
            for (int na = 0; na < *na_atcell; na++)
              for (size_t i = 0; i < ra; i++) {
                rhotrur[i + ra * ((ns - 1) + (size_t)nspv * na)] = t1[i + ra * ((ns1 - 1) + (size_t)nspv1 * na)] * fctr;
                rhosmtr[i + ra * ((ns - 1) + (size_t)nspv * na)] = u1[i + ra * ((ns1 - 1) + (size_t)nspv1 * na)] * fctr;
              }
          }
          free(s1); free(t1); free(u1);
        }
      }
    done:
      fclose(f);
      if (rc) return rc;
    }
  }
  if (!histyn) *lbrydn = 0;
  return 0;
}

/* ------------------------------------------------------------------ rw_wfpot8f.F90:305-349 read_vht */
int orc_read_vht(int nrIn, const char *chdir, int nevhist, int ncpx, int ncpy, int ncpz, int nmesh,
                 double *vh_dense) {
  const size_t nd = (size_t)ncpx * nmesh * ncpy * nmesh * ncpz * nmesh;
  int histyn = nevhist == 2 || nevhist == 3 || nevhist == -1;
  const char *E = "error found in vht.*";
  if (histyn) {
    char path[1100];
    int rc = orc_ckpt_path(chdir, "vht", nrIn < 0 ? 0 : nrIn, -1, path, sizeof path);
    if (rc) return rc;
    FILE *f = fopen(path, "rb");
    histyn = f != NULL;
    if (histyn) {
      int ver = 0, d[4];
      rc = get_rec(f, &ver, 4, E);
      if (!rc && ver != VHT_VERSION) rc = orc_set_error(1, "wrong vht version");
      if (!rc) rc = get_rec(f, d, sizeof d, E);
      if (!rc) rc = get_rec(f, NULL, 0, E);
      if (!rc && (d[0] != ncpx || d[1] != ncpy || d[2] != ncpz || d[3] != nmesh)) {
        static char msg[1400];
        snprintf(msg, sizeof msg, "read_vht: unexpected array dimensions in potential file %s", path);
        rc = orc_set_error(1, msg);
      }
      if (!rc) rc = get_rec(f, vh_dense, 8 * nd, E);
      fclose(f);
      if (rc) return rc;
    }
  }
  if (!histyn) memset(vh_dense, 0, 8 * nd);
  return 0;
}

/* ------------------------------------------------------------------ rw_wfpot8f.F90:402-505 write_rhovht */
int orc_write_rhovht(int nrIn, const char *chdir, int nrho, int nele, int npre, int nevhist, int nspv,
                     int ncpx, int ncpy, int ncpz, int nmesh, int nradmx, int npoint, int num_atcell,
                     int nrxIn, int nryIn, int nrzIn, int natom, int key_natpri_in, const int *natpri,
                     const double *rhosmt_i, const double *rhotrur_i, const double *rhosmtr_i,
                     const double *rhosmt_o, const double *rhotrur_o, const double *rhosmtr_o,
                     const double *vh_dense) {
  const int nr = nrIn < 0 ? 0 : nrIn;
  const int rk[3] = {nrxIn < 0 ? 0 : nrxIn, nryIn < 0 ? 0 : nryIn, nrzIn < 0 ? 0 : nrzIn};
  const size_t n = (size_t)ncpx * ncpy * ncpz, ra = (size_t)nradmx * npoint;
  char path[1100];
  int rc = 0;
  if (nevhist == 1 || nevhist == 3) {
    for (int irho = 1; irho <= nrho && !rc; irho++) {
      if ((rc = orc_ckpt_path(chdir, irho == 1 ? "rho" : "r_o", nr, -1, path, sizeof path))) return rc;
      FILE *f = fopen(path, "wb");
      if (!f) return orc_set_error(1, "cannot open rho.* file(s)");
      if (npre == 2) {
        fclose(f);
        remove(path);
        continue;
      }
      int *rec = (int *)calloc((size_t)num_atcell + 1, sizeof(int));
      int na_atcell = 0;
      for (int na = 1; na <= natom && !rc; na++)
        if (natpri[na - 1] == key_natpri_in) {
          na_atcell++;
          if (na_atcell > num_atcell) rc = orc_set_error(1, "write_rhovht: na_atcell > num_atcell");
          else rec[na_atcell] = na;
        }
      rec[0] = na_atcell;
      const int num_atcell1 = na_atcell;  /* l_samedim = .false. (:420) */
      const int ver = RHO_VERSION, d[7] = {nspv, ncpx, ncpy, ncpz, nradmx, npoint, num_atcell1};
      if (!rc) rc = put1(f, &ver, 4);
      if (!rc) rc = put1(f, d, sizeof d);
      if (!rc) rc = put1(f, rk, sizeof rk);
      if (!rc) rc = put1(f, rec, 4 * ((size_t)num_atcell + 1));
      const double *s = irho == 1 ? rhosmt_i : rhosmt_o, *t = irho == 1 ? rhotrur_i : rhotrur_o;
      const double *u = irho == 1 ? rhosmtr_i : rhosmtr_o;
      if (!rc) rc = put1(f, s, 8 * n * nspv);
      if (!rc) rc = put1(f, t, 8 * ra * nspv * num_atcell1);
      if (!rc) rc = put1(f, u, 8 * ra * nspv * num_atcell1);
      free(rec);
      fclose(f);
    }
  }
  if (!rc && (nevhist == 2 || nevhist == 3)) {
    if ((rc = orc_ckpt_path(chdir, "vht", nr, -1, path, sizeof path))) return rc;
    FILE *f = fopen(path, "wb");
    if (!f) return orc_set_error(1, "cannot open vht.* file(s)");
    if (npre == 2) {
      fclose(f);
      remove(path);
    } else {
      const int ver = VHT_VERSION, d[4] = {ncpx, ncpy, ncpz, nmesh};
      rc = put1(f, &ver, 4);
      if (!rc) rc = put1(f, d, sizeof d);
      if (!rc) rc = put1(f, rk, sizeof rk);
      if (!rc) rc = put1(f, vh_dense, 8 * n * (size_t)nmesh * nmesh * nmesh);
      fclose(f);
    }
  }
  if (!rc && nele == 1) {
    if ((rc = orc_ckpt_path(chdir, "ele", nr, -1, path, sizeof path))) return rc;
    FILE *f = fopen(path, "wb");
    if (!f) return orc_set_error(1, "cannot open ele.* file(s)");
    rc = put1(f, rhosmt_o, 8 * n * nspv);
    fclose(f);
  }
  return rc;
}

/* ------------------------------------------------------------------ rw_wfpot8f.F90:506-532 */
int orc_filenumbers_rhovht(const char *chdir, int *nrrho, int *nrpot) {
  char path[1100];
  const char *names[2] = {"rho", "vht"};
  int *outs[2] = {nrrho, nrpot};
  for (int k = 0; k < 2; k++) {
    int nr = 0;
    for (;;) {
      int rc = orc_ckpt_path(chdir, names[k], nr, -1, path, sizeof path);
      if (rc) return rc;
      FILE *f = fopen(path, "rb");
      if (!f) break;
      fclose(f);
      nr++;
    }
    *outs[k] = nr;
  }
  return 0;
}
