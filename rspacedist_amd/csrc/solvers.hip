// solvers.hip — the Kohn-Sham hot path on device pointers (libkukan_hip.so, gfx950).
//
// Each routine restates the reference routine named in its comment; the entry points in
// abi.hip resolve the caller's arrays (host, bound or device) and call these.
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>

#include "gemm_mixed.hpp"
#include "kukan_internal.hpp"
#include "solvers.hpp"

namespace kukan {
namespace {

inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

// ----------------------------------------------------------------------------- scalar kernels
__global__ void k_rsqrt_norm(int nb, const double *vecnor, double *tmp, int *flag) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double v = vecnor[b];
  if (!(v > 0.0)) {
    atomicOr(flag, 1);
    tmp[b] = 0.0;
  } else {
    tmp[b] = 1.0 / sqrt(v);
  }
}

// kscg8f.F90:1197-1230 (kscg_r_05) for each band; inactive bands keep alph=1, beta=0.
__global__ void k_kscg05(int nb, const double *sval, const double *pap, const double *sapre,
                         const int *active, double *alph, double *beta) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double a = 1.0, be = 0.0;
  if (!active || active[b]) {
    const double eps = 1.0e-14, sv = sval[b], pp = pap[b], sa = sapre[b];
    const double gamma = (sv + pp) * 0.5 - sqrt((sv - pp) * (sv - pp) * 0.25 + sa * sa);
    if (fabs((sv - gamma) / gamma) > eps) {
      a = sa / sqrt(sa * sa + (sv - gamma) * (sv - gamma));
      be = -(sv - gamma) / sqrt(sa * sa + (sv - gamma) * (sv - gamma));
    }
    if (a < eps) {
      a = -a;
      be = -be;
    }
  }
  alph[b] = a;
  beta[b] = be;
}

// f[b] <- 1 / sqrt(f[b]); a non-positive norm raises the flag
__global__ void k_rsqrt(int nb, double *f, int *flag) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  if (f[b] > 0.0) {
    f[b] = 1.0 / sqrt(f[b]);
  } else {
    f[b] = 1.0;
    *flag = 1;
  }
}

// zero the strictly lower triangle of an M x M column-major matrix
__global__ void k_zero_lower(int M, double *A) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)M * M) return;
  const int i = (int)(t % M), j = (int)(t / M);
  if (i > j) A[t] = 0.0;
}

// t[b] = reset ? 0 : num[b]/den[b]
__global__ void k_ratio(int nb, const double *num, const double *den, int reset, double *t) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  t[b] = reset ? 0.0 : num[b] / den[b];
}

// block CG direction scalars: d = <S psi|P>, then normalisation of P - d psi.
//   q = <P|S P> - 2 d <psi|S P>... with S psi, psi orthonormal-ish we use the exact
//   expression n2 = <P|SP> - 2 d e + d^2 g, e = <Spsi|P>=d, g = <psi|Spsi>.
__global__ void k_pnorm(int nb, const double *d, const double *pp, const double *g,
                        double *fac, double *mdfac, int *flag) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double n2 = pp[b] - 2.0 * d[b] * d[b] + d[b] * d[b] * g[b];
  if (!(n2 > 0.0)) {
    atomicOr(flag, 1);
    fac[b] = 0.0;
    mdfac[b] = 0.0;
    return;
  }
  const double f = 1.0 / sqrt(n2);
  fac[b] = f;
  mdfac[b] = -d[b] * f;
}

// psi <- al psi + be p with p = fac P + mdf psi, as psi <- (al + be mdf) psi + (be fac) P
__global__ void k_comb_coef(int nb, const double *al, const double *be, const double *fac,
                            const double *mdf, double *a2, double *c2) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  a2[b] = al[b] + be[b] * mdf[b];
  c2[b] = be[b] * fac[b];
}

// flag = 1 if some dij(:, :, s, a) is not exactly symmetric (count = nprjmx^2 nums natom)
__global__ void k_dij_asym(const double *D, int np, long count, int *flag) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= count) return;
  const long m = i / ((long)np * np);
  const int r = (int)(i % ((long)np * np)), a = r % np, b = r / np;
  const double *Dm = D + m * np * np;
  if (Dm[a + (long)np * b] != Dm[b + (long)np * a]) *flag = 1;
}

// symmetric form of the combination dots (p = fac x + mdf y, H symmetric, all times dV):
// sap = <fac x|H y> + mdf <y|H y>, then pap = <fac x|H p> + mdf sap
__global__ void k_comb_sym(int nb, const double *mdf, const double *eh, double *pap, double *sap) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double sy = sap[b] + mdf[b] * eh[b];
  sap[b] = sy;
  pap[b] += mdf[b] * sy;
}

// eps of psi_new = al psi + be p for a symmetric H (all times dV):
// al^2 <psi|H psi> + 2 al be <psi|H p> + be^2 <p|H p>
__global__ void k_eps_next(int nb, const double *al, const double *be, const double *eh,
                           const double *sap, const double *pap, double *eps) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double x = al[b], y = be[b];
  eps[b] = x * x * eh[b] + 2.0 * x * y * sap[b] + y * y * pap[b];
}

// ksdiis8f.F90:285-600 per band, nitermax = 2 (one trial step), expressed on the basis
// (phi0, Kr): returns coefficients (ca, cb) of the new band and its residual measure.
// Inputs per band (already multiplied by dV): e = eps, hkk=<Kr|HKr>, h0k=<phi0|HKr>,
// hk0=<Kr|Hphi0>, s00=<phi0|Sphi0>, sk0=<Kr|Sphi0>, skk=<Kr|SKr>, r00=<R0|R0>, r0q=<R0|Q>,
// qq=<Q|Q> with R0 = Hphi0 - e Sphi0, Q = HKr - e SKr.
struct DiisIn {
  double h00, hkk, h0k, hk0, s00, sk0, skk, r00, r0q, qq;
};
__device__ void sym2_eig(double a, double b, double c, double *w, double *v) {
  // [[a, b],[b, c]]: ascending eigenvalues, eigenvectors canonical (largest |.| positive)
  const double tr = 0.5 * (a + c), df = 0.5 * (a - c);
  const double r = hypot(df, b);
  w[0] = tr - r;
  w[1] = tr + r;
  for (int k = 0; k < 2; k++) {
    double x, y;
    // (A - w I) v = 0
    if (fabs(a - w[k]) + fabs(b) >= fabs(c - w[k]) + fabs(b)) {
      x = -b;
      y = a - w[k];
    } else {
      x = c - w[k];
      y = -b;
    }
    double nn = hypot(x, y);
    if (nn == 0.0) {
      x = k == 0 ? 1.0 : 0.0;
      y = k == 0 ? 0.0 : 1.0;
      if (b == 0.0 && a > c) { x = 1.0 - x; y = 1.0 - y; }
      nn = 1.0;
    }
    x /= nn;
    y /= nn;
    const bool flip = (fabs(x) >= fabs(y)) ? (x < 0.0) : (y < 0.0);
    if (flip) { x = -x; y = -y; }
    v[2 * k] = x;
    v[2 * k + 1] = y;
  }
}

__global__ void k_diis_step(int nb, const double *dots /*10 per band*/, double dv, double a1,
                            double lmin, double lmax, double eps_eig, double *sval_io,
                            double *lambda, double *ca, double *cb, double *resid) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double *d = dots + 10 * b;
  const double e = d[0] * dv;  // eps = dV <phi0|H phi0>
  sval_io[b] = e;
  // line search (ksdiis8f.F90:296-320): phi1' = phi0 + a1 Kr
  const double s1 = d[3], s2 = d[5];
  const double s3 = d[0] + a1 * (d[2] + d[3]) + a1 * a1 * d[1];
  const double s4 = d[4] + 2.0 * a1 * d[5] + a1 * a1 * d[6];
  const double aaa = 2.0 * (s1 - e * s2) * dv;
  const double vr = s3 / s4;
  const double bbb = (vr - e - aaa * a1) / (a1 * a1);
  double lam = 0.5 * aaa / bbb;
  if (lam > lmax) lam = lmax;
  if (lam < lmin) lam = lmin;
  lambda[b] = lam;
  // rsmat/rrmat of {phi0, phi1 = phi0 + lam Kr} (:375-408), then phi1 normalised
  double S00 = d[4] * dv, S01 = (d[4] + lam * d[5]) * dv,
         S11 = (d[4] + 2.0 * lam * d[5] + lam * lam * d[6]) * dv;
  double R00 = d[7] * dv, R01 = (d[7] + lam * d[8]) * dv,
         R11 = (d[7] + 2.0 * lam * d[8] + lam * lam * d[9]) * dv;
  const double tmp = 1.0 / sqrt(S11);
  S01 *= tmp; R01 *= tmp; S11 *= tmp * tmp; R11 *= tmp * tmp;
  // overlap eigenvectors, descending (:410-420)
  double w[2], v[4], rv[4], eg[2];
  sym2_eig(S00, S01, S11, w, v);
  rv[0] = v[2]; rv[1] = v[3]; eg[0] = w[1];  // column 1 = largest
  rv[2] = v[0]; rv[3] = v[1]; eg[1] = w[0];
  int nmat = 0, k = 0;
  while (k < 2) {
    nmat++;
    k++;
    if (eg[nmat - 1] < eps_eig) {
      // columns nmat..iiter-1 shift left (iiter = 1: nothing moves)
      nmat--;
    }
  }
  double c0, c1;
  double res = resid[b];
  if (nmat > 1) {
    for (int i = 0; i < nmat; i++) {
      const double f = 1.0 / sqrt(eg[i]);
      rv[2 * i] *= f;
      rv[2 * i + 1] *= f;
    }
    // rrmat2 = rv^T R rv
    const double R[4] = {R00, R01, R01, R11};
    double T[4], M2[4];
    for (int kk = 0; kk < 2; kk++)
      for (int j = 0; j < 2; j++) T[kk + 2 * j] = R[kk + 0] * rv[0 + 2 * j] + R[kk + 2] * rv[1 + 2 * j];
    for (int kk = 0; kk < 2; kk++)
      for (int j = 0; j < 2; j++)
        M2[kk + 2 * j] = rv[0 + 2 * kk] * T[0 + 2 * j] + rv[1 + 2 * kk] * T[1 + 2 * j];
    double w2[2], v2[4];
    sym2_eig(M2[0], M2[2], M2[3], w2, v2);  // dsyev('v','u'): upper triangle
    // descending: column nmat (=2) is the smallest eigenvalue
    res = w2[0];
    const double x = v2[0], y = v2[1];
    c0 = rv[0] * x + rv[2] * y;
    c1 = rv[1] * x + rv[3] * y;
    const double S[4] = {S00, S01, S01, S11};
    const double q = c0 * (S[0] * c0 + S[2] * c1) + c1 * (S[1] * c0 + S[3] * c1);
    const double f = 1.0 / sqrt(q);
    c0 *= f;
    c1 *= f;
  } else {
    c0 = 0.0;
    c1 = 1.0;
  }
  if (c0 < 0.0) { c0 = -c0; c1 = -c1; }
  resid[b] = res;
  // new = c0 phi0 + c1 tmp (phi0 + lam Kr)
  ca[b] = c0 + c1 * tmp;
  cb[b] = c1 * tmp * lam;
}

// ten DIIS dots per band in one pass: see DiisIn ordering in k_diis_step
__global__ __launch_bounds__(256) void k_diis_dots(long n, const double *phi, const double *hphi,
                                                   const double *sphi, const double *kr,
                                                   const double *hkr, const double *skr, long ld,
                                                   const double *sval_dv /*eps per band*/,
                                                   double *part, int nparts) {
  __shared__ double red[10][4];
  const int b = blockIdx.y;
  const long chunk = (n + nparts - 1) / nparts;
  const long i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  const long o = (long)b * ld;
  const double e = sval_dv[b];
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double p = phi[o + i], hp = hphi[o + i], sp = sphi[o + i];
    const double k = kr[o + i], hk = hkr[o + i], sk = skr[o + i];
    const double r0 = hp - e * sp, q = hk - e * sk;
    acc[0] += p * hp;
    acc[1] += k * hk;
    acc[2] += p * hk;
    acc[3] += k * hp;
    acc[4] += p * sp;
    acc[5] += k * sp;
    acc[6] += k * sk;
    acc[7] += r0 * r0;
    acc[8] += r0 * q;
    acc[9] += q * q;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j = 0; j < 10; j++) {
    double v = acc[j];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[j][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    const int j = threadIdx.x;
    part[((long)b * 10 + j) * nparts + blockIdx.x] = ((red[j][0] + red[j][1]) + red[j][2]) + red[j][3];
  }
}
__global__ __launch_bounds__(256) void k_diis_reduce(const double *part, int nparts, double *dots) {
  __shared__ double red[4];
  const int bj = blockIdx.x;  // b*10 + j
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) s += part[(long)bj * nparts + i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) dots[bj] = ((red[0] + red[1]) + red[2]) + red[3];
}
// eps per band (dV <phi0|Hphi0>) is needed inside k_diis_dots; computed by a dot first.

// FD first derivative along axis (force8f.F90:408-1010): out = sum_m w_m (psi(+m)-psi(-m))/h
__global__ void k_fd1(int nx, int ny, int nz, int px, int py, int pz, int axis, int nf,
                      const double *w, double hinv, const double *in, long ld, double *out,
                      long ldo) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long n = (long)nx * ny * nz;
  const int b = blockIdx.y;
  if (t >= n) return;
  const int x = (int)(t % nx), y = (int)((t / nx) % ny), z = (int)(t / ((long)nx * ny));
  const double *v = in + (long)b * ld;
  auto at = [&](int dxx, int dyy, int dzz) -> double {
    int X = x + dxx, Y = y + dyy, Z = z + dzz;
    if (X < 0 || X >= nx) { if (!px) return 0.0; X = X < 0 ? X + nx : X - nx; }
    if (Y < 0 || Y >= ny) { if (!py) return 0.0; Y = Y < 0 ? Y + ny : Y - ny; }
    if (Z < 0 || Z >= nz) { if (!pz) return 0.0; Z = Z < 0 ? Z + nz : Z - nz; }
    return v[(long)X + (long)nx * ((long)Y + (long)ny * Z)];
  };
  const int sx = axis == 0, sy = axis == 1, sz = axis == 2;
  double s = 0.0;
  for (int m = nf; m >= 1; m--) s += (w[m] * hinv) * at(sx * m, sy * m, sz * m);
  for (int m = 1; m <= nf; m++) s -= (w[m] * hinv) * at(-sx * m, -sy * m, -sz * m);
  out[(long)b * ldo + t] = s;
}

// nonlocal band force accumulation (force8f.F90:149-165), one thread per active atom,
// bands in order; only natpri == key_natpri_in atoms contribute.
__global__ void k_force_acc(const int *act, int nact, const int *is_in, int nprjmx, int nums,
                            int ns, const double *dij, const double *sss, const double *rp,
                            const double *rd, long ld_rp, int nb, const double *fnele,
                            const double *sval, double *fa) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nact || !is_in[a]) return;
  const int slot = act[5 * a], iprj = act[5 * a + 2], na = act[5 * a + 3], ispe = act[5 * a + 4];
  const double *D = dij + (long)nprjmx * nprjmx * ((ns - 1) + (long)nums * na);
  const double *S = sss + (long)nprjmx * nprjmx * ispe;
  double f = fa[na];
  for (int l = 0; l < nb; l++) {
    const double *p = rp + (long)l * ld_rp + (long)nprjmx * slot;
    const double *q = rd + (long)l * ld_rp + (long)nprjmx * slot;
    const double fl = fnele[l], ev = sval[l];
    for (int j = 0; j < iprj; j++)
      for (int i = 0; i < iprj; i++)
        f = f - 2.0 * fl * (p[i] * q[j]) * (D[i + (long)nprjmx * j] - ev * S[i + (long)nprjmx * j]);
  }
  fa[na] = f;
}

// rho(:,s) += sum_l f_l psi_l^2 over a band chunk, band order (scf_charge8f.F90:303-318)
__global__ void k_charge(long n, int nb, const double *psi, long ld, const double *f, double *rho) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r = rho[i];
  for (int l = 0; l < nb; l++) {
    const double w = psi[(long)l * ld + i];
    r = r + f[l] * w * w;
  }
  rho[i] = r;
}

// x *= alpha
__global__ void k_scal(long n, double alpha, double *x) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) x[i] *= alpha;
}

// Nonlocal part of dE/dR for E = sum_l f_l sum_ij rp_i(l) D_ij rp_j(l): work-group (active
// atom a, direction d) reduces over bands l of
//   f_l sum_ij (drp_d,i(l) D_ij rp_j(l) + rp_i(l) D_ij drp_d,j(l))
// in a fixed order.  drp_d = <d p / d R_d | psi> from the derivative tables.
__global__ __launch_bounds__(256) void k_adj_force(const int *act, int nprjmx, int nums,
                                                   const double *dij, const double *rp,
                                                   const double *drp, long n1, int M,
                                                   const double *f, const int *is_in,
                                                   int natom, double *dedr) {
  __shared__ double red[256];
  const int a = blockIdx.x, d = blockIdx.y;
  const int slot = act[5 * a], iprj = act[5 * a + 2], na = act[5 * a + 3];
  const double *D = dij + (long)nprjmx * nprjmx * ((long)nums * na);
  const double *dr = drp + (long)d * n1 * M;
  double acc = 0.0;
  for (int l = threadIdx.x; l < M; l += 256) {
    const double *r = rp + (long)l * n1 + (long)nprjmx * slot;
    const double *q = dr + (long)l * n1 + (long)nprjmx * slot;
    double s = 0.0;
    for (int j = 0; j < iprj; j++)
      for (int i = 0; i < iprj; i++) s += (q[i] * r[j] + r[i] * q[j]) * D[i + (long)nprjmx * j];
    acc += f[l] * s;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && is_in[a]) dedr[(long)d * natom + na] = red[0];
}

// Hartree: b = 2 pi rho; boundary planes of isolated directions (scf_hartree8f.F90:88-595)
__global__ void k_hartree_b(long n, const double *rho, double *b) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = 2.0 * 3.14159265358979323846 * rho[i];
}
__global__ void k_hartree_bound(int nx, int ny, int nz, int nfh, int axis, const double *w,
                                const double *bound, double *b) {
  // one thread per face point; lower and upper faces
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long d1, d2;
  if (axis == 0) { d1 = ny; d2 = nz; } else if (axis == 1) { d1 = nx; d2 = nz; } else { d1 = nx; d2 = ny; }
  if (t >= d1 * d2) return;
  const int u = (int)(t % d1), v = (int)(t / d1);
  const int nb2 = 2 * nfh;
  for (int m = 1; m <= nfh; m++)
    for (int q = 0; q <= nfh - m; q++) {
      const int mm = m + q;
      long ilo, ihi, olo, ohi;
      if (axis == 2) {
        ilo = u + (long)nx * (v + (long)ny * (1 - m + nfh - 1));
        ihi = u + (long)nx * (v + (long)ny * (m + nfh - 1));
        olo = u + (long)nx * (v + (long)ny * q);
        ohi = u + (long)nx * (v + (long)ny * (nz - 1 - q));
      } else if (axis == 1) {
        ilo = u + (long)nx * ((1 - m + nfh - 1) + (long)nb2 * v);
        ihi = u + (long)nx * ((m + nfh - 1) + (long)nb2 * v);
        olo = u + (long)nx * (q + (long)ny * v);
        ohi = u + (long)nx * ((ny - 1 - q) + (long)ny * v);
      } else {
        ilo = (1 - m + nfh - 1) + (long)nb2 * (u + (long)ny * v);
        ihi = (m + nfh - 1) + (long)nb2 * (u + (long)ny * v);
        olo = q + (long)nx * (u + (long)ny * v);
        ohi = (nx - 1 - q) + (long)nx * (u + (long)ny * v);
      }
      b[olo] -= w[mm] * bound[ilo];
      b[ohi] -= w[mm] * bound[ihi];
    }
}
struct Nine {
  double v[9];
};
__global__ void k_set9(Nine a, double *out) {
  if (threadIdx.x < 9) out[threadIdx.x] = a.v[threadIdx.x];
}

__global__ void k_sub_scalar(long n, const double *s, double f, double *x) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] - s[0] * f;
}
// Hartree CG updates (scf_hartree8f.F90:849-891); scalars on the device: sc[0]=r1r, sc[1]=pap
// Device-side Hartree CG loop state: st[0] = done, st[1] = current iteration (1-based).
// Every kernel of an iteration returns at once when done is set, so a block of iterations
// can be launched (or replayed as a graph) past convergence without changing the result.
// An iteration is four launches: A p (stencil, with <p|A p> partials), the update (each
// work-group folds the stencil partials itself), the direction (each work-group folds the
// update's <r|r> partials itself) and the stopping rule of scf_hartree8f.F90.
__global__ void k_hcg_reset(int *st) {
  st[0] = 0;
  st[1] = 1;
}
// Sum of np partials by a 256-thread work-group in a fixed order (strided per thread, then
// wave shuffles, then the 4 wave sums in order); every thread gets the result.
__device__ inline double fold_parts256(const double *part, int np, double *red4) {
  double s = 0.0;
  for (int c = threadIdx.x; c < np; c += 256) s += part[c];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = s;
  __syncthreads();
  const double tot = ((red4[0] + red4[1]) + red4[2]) + red4[3];
  __syncthreads();
  return tot;
}
__global__ void k_hcg_update(long n, const double *sc, double *vh, const double *p, const double *ap,
                             double *avhs, const double *b, double *r, const double *spart,
                             int nsp, double *part, int nparts, const int *st) {
  __shared__ double red[4];
  if (st[0]) return;
  const double alph = sc[0] / fold_parts256(spart, nsp, red);  // r1r / pap
  const long chunk = (n + nparts - 1) / nparts;
  const long i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  double s = 0.0;
  for (long i = i0 + threadIdx.x; i < i1; i += 256) {
    vh[i] = vh[i] + alph * p[i];
    avhs[i] = avhs[i] + alph * ap[i];
    const double ri = b[i] - avhs[i];
    r[i] = ri;
    s += ri * ri;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}
__global__ void k_hcg_dir(long n, double *sc, const double *upart, int nup, int ncgres,
                          const double *r, double *p, const int *st) {
  __shared__ double red[4];
  if (st[0]) return;
  const double r2r = fold_parts256(upart, nup, red);
  const double beta = (st[1] % ncgres == 0) ? 0.0 : r2r / sc[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[2] = r2r;
  const long stride = (long)gridDim.x * 256;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += stride) p[i] = r[i] + beta * p[i];
}
// stopping rule (ncgmin, ncgmax, dp = sqrt(r.r dV / Omega) <= epsvh); else next iteration
__global__ void k_hcg_ctl(double *sc, double dvo, double eps, int ncgmin, int ncgmax, int *st) {
  if (st[0]) return;
  const int loop = st[1];
  const double dp = sqrt(sc[2] * dvo);
  if (loop >= ncgmin && (loop >= ncgmax || dp <= eps)) {
    st[0] = 1;
    return;
  }
  st[1] = loop + 1;
  sc[0] = sc[2];
}


// ---------------------------------------------------------------- persistent Hartree CG
// The whole CG loop of scf_hartree8f.F90:611-1143 in one cooperative launch (opt-in, param
// hartree_persistent 1).  Work-group g owns a 16 x 16 (x, y) patch times 16 z-planes; each
// thread keeps its z-column's V, A V, A p and p in registers for the whole solve, so an
// iteration moves only p and b through memory: each thread stores its new p, and the stencil
// reads the patch's x / y halo tiles (8 planes at a time, by LDS-DMA) and each column's z-halo
// from the global copy.  Per iteration one grid barrier (p complete) and two all-gathered sums
// (<p|A p> and <r|r>: every work-group reads all work-groups' partials and folds them in the
// same order, so all of them take the same alpha, beta and stopping decision (ncgmin, ncgmax,
// dp <= epsvh, restart every ncgres) and leave the loop together).
// Grid barrier state, cleared by k_hcgp_init before each solve: `err` (a wait ran out), then
// one arrival word per work-group (the last generation it reached).
struct HcgBar {
  unsigned err, pad[63];
  unsigned arrive[1];  // gridDim.x words
};
struct HcgSlot;
__global__ void k_hcgp_init(HcgBar *bar, int nblocks, unsigned *slot_gens, int nslots, int *state) {
  for (int i = threadIdx.x; i < nblocks; i += blockDim.x) bar->arrive[i] = 0;
  for (int i = threadIdx.x; i < nslots; i += blockDim.x) slot_gens[4 * i + 2] = 0;  // HcgSlot::gen
  if (threadIdx.x == 0) {
    bar->err = 0;
    state[0] = 0;
    state[1] = 0;
  }
}
constexpr int HCGP_T = 16, HCGP_P = 16;  // patch edge in x and y; z-points per thread
constexpr long HCGP_SPINS = 1L << 22;    // bound of every wait (seconds; a real wait is microseconds)
constexpr int HCGP_AUX_SC1 = 16;         // buffer cache policy: sc1 (agent-coherent), as the atomics
// Data that crosses work-groups inside the solve (p, the partial sums) is written and read only
// with agent-scope atomics (coherent across the XCDs' L2 caches, no cache-wide write-back or
// invalidate); the barrier orders them by waiting for each wave's stores to complete before the
// work-group arrives, and compiler-only fences keep the compiler from moving them across it.
__device__ inline void hcgp_st(double *q, double v) {
  __hip_atomic_store(q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double hcgp_ld(const double *q) {
  return __hip_atomic_load(const_cast<double *>(q), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned hcgp_ldu(unsigned *q) {
  return __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void hcgp_stu(unsigned *q, unsigned v) {
  __hip_atomic_store(q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// wait until *w holds g1 (bounded; a wait that runs out marks the barrier failed)
__device__ inline bool hcgp_wait(unsigned *w, unsigned g1, HcgBar *bar) {
  long spins = 0;
  while (hcgp_ldu(w) != g1) {
    ++spins;
    if (((spins & 255) == 0 && hcgp_ldu(&bar->err) != 0u) || spins > HCGP_SPINS) {
      hcgp_stu(&bar->err, 1u);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}
// Every work-group arrives before any leaves: each publishes the new generation in its own
// arrival word and every work-group's threads wait for all words in parallel.  Arrivals are
// independent stores, so nothing serialises on one address.  The launch is cooperative (all
// work-groups resident); a wait that runs out marks the barrier failed, and every work-group
// then leaves the loop.
__device__ inline bool hcgp_barrier(HcgBar *bar, unsigned nblocks, unsigned &gen, int *ok_sh) {
  const unsigned g1 = ++gen;
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores have completed
  __syncthreads();
  if (threadIdx.x == 0) {
    hcgp_stu(&bar->arrive[blockIdx.x], g1);
    *ok_sh = 1;
  }
  __syncthreads();
  for (unsigned i = threadIdx.x; i < nblocks; i += blockDim.x)
    if (!hcgp_wait(&bar->arrive[i], g1, bar)) *ok_sh = 0;
  __syncthreads();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return *ok_sh != 0 && hcgp_ldu(&bar->err) == 0u;
}
// Sum over all work-groups of one value each, every work-group getting the same result: each
// publishes its value and then the generation in its slot (the value's store completes first),
// and every work-group reads all slots once their generation is current and folds them in slot
// order (strided per thread, wave shuffles, the 4 wave sums in order).  ok is cleared when a
// wait runs out.
struct HcgSlot {
  double v;
  unsigned gen, pad;
};
__device__ inline double hcgp_allsum(double mine, HcgSlot *slot, HcgBar *bar, unsigned nblocks,
                                     unsigned &gen, double *red4, int *ok_sh) {
  const unsigned g1 = ++gen;
  for (int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off, 64);
  if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    hcgp_st(&slot[blockIdx.x].v, ((red4[0] + red4[1]) + red4[2]) + red4[3]);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // the value is stored before the generation
    hcgp_stu(&slot[blockIdx.x].gen, g1);
    *ok_sh = 1;
  }
  __syncthreads();
  for (unsigned i = threadIdx.x; i < nblocks; i += blockDim.x)  // all of this thread's slots
    if (!hcgp_wait(&slot[i].gen, g1, bar)) *ok_sh = 0;          // current first, then read
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  double s = 0.0;
  for (unsigned i = threadIdx.x; i < nblocks; i += blockDim.x) s += hcgp_ld(&slot[i].v);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __syncthreads();  // every thread has read red4 above
  if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = s;
  __syncthreads();
  const double tot = ((red4[0] + red4[1]) + red4[2]) + red4[3];
  __syncthreads();
  return tot;
}
struct HPArgs {
  int nx, ny, nz, px, py, pz, bx, by;
  double cx[7], cy[7], cz[7], c0;
  double *vh, *p;
  HcgSlot *slot1, *slot2;  // <p|A p> and <r|r> partials, one slot per work-group
  const double *avhs, *b, *sc;  // sc[0]: <r|r> of the start residual
  double dvo, eps;
  int ncgmin, ncgmax, ncgres;
  int *state;  // [0]: 1 done, 2 a grid barrier failed; [1]: iterations
  HcgBar *bar;
};
__device__ inline int hcgp_wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }  // |overshoot| < n
// p at (X, Y, z) with the grid's boundary rule: periodic directions wrap, isolated ones are 0
__device__ inline double hcgp_p(const HPArgs &a, int X, int Y, int z) {
  if (X < 0 || X >= a.nx) { if (!a.px) return 0.0; X = hcgp_wrap(X, a.nx); }
  if (Y < 0 || Y >= a.ny) { if (!a.py) return 0.0; Y = hcgp_wrap(Y, a.ny); }
  if (z < 0 || z >= a.nz) { if (!a.pz) return 0.0; z = hcgp_wrap(z, a.nz); }
  return hcgp_ld(a.p + (long)X + (long)a.nx * ((long)Y + (long)a.ny * z));
}
// nx, ny, nz are multiples of 16 (the host checks): every thread owns 16 grid points.  The
// stencil stages the x / y halo of 8 planes at a time in LDS (one load latency per 8 planes)
// and the z-halo of every column with the first 8.
template <int H>
__global__ __launch_bounds__(256, 2) void k_hcg_persistent(HPArgs a) {
  constexpr int T = HCGP_T, P = HCGP_P, W = T + 2 * H, PG = 8;  // planes per staging phase
  __shared__ double tile[PG][W * W];
  __shared__ double zh[2 * H][256];   // each column's p below (q < H) and above the segment
  __shared__ double cf[3 * (H + 1)];  // cx, cy, cz
  __shared__ double red4[4];
  __shared__ int ok_sh;
  const int t = threadIdx.x, tx = t % T, ty = t / T;
  const unsigned G = gridDim.x;
  const int gi = blockIdx.x;
  const int x0 = (gi % a.bx) * T, y0 = ((gi / a.bx) % a.by) * T, z0 = (gi / (a.bx * a.by)) * P;
  const int x = x0 + tx, y = y0 + ty;
  const long plane = (long)a.nx * a.ny;
  double *pcol = a.p + (long)x + (long)a.nx * y;  // this thread's column of p
  const long c0 = plane * z0;                     // its segment's first point, from pcol
  if (t <= H) {
    cf[t] = a.cx[t];
    cf[H + 1 + t] = a.cy[t];
    cf[2 * (H + 1) + t] = a.cz[t];
  }
  double vh[P], av[P], ap[P], e[P];  // e[j]: p at z0 + j (b is read when used)
#pragma unroll
  for (int j = 0; j < P; j++) {
    const long i = (long)x + (long)a.nx * y + plane * (z0 + j);
    vh[j] = a.vh[i];
    av[j] = a.avhs[i];
    e[j] = a.p[i];
    ap[j] = 0.0;
  }
  double r1r = a.sc[0];
  unsigned gen = 0;
  int loop = 1, status = 1;
  for (;;) {
    if (!hcgp_barrier(a.bar, G, gen, &ok_sh)) { status = 2; break; }  // p complete
    // addresses are formed anew each iteration: hoisting them out of the loop would keep
    // every plane's halo addresses live at once
    HPArgs ai = a;
    asm volatile("" : "+s"(ai.p));
    const double *pc = ai.p + (long)x + (long)a.nx * y;
    const __amdgpu_buffer_rsrc_t prs =
        __builtin_amdgcn_make_buffer_rsrc((void *)ai.p, (short)0, (int)(plane * a.nz * 8), 0x00020000);
    for (int q = 0; q < 2 * H; q++) {  // z-planes of the halo: wrapped, or 0 outside an isolated cell
      int zq = q < H ? z0 - H + q : z0 + P + q - H;
      const bool out = zq < 0 || zq >= a.nz;
      zq = hcgp_wrap(zq, a.nz);
      zh[q][t] = (out && !a.pz) ? 0.0 : hcgp_ld(pc + plane * zq);
    }
    double pap = 0.0;
#pragma unroll
    for (int j0 = 0; j0 < P; j0 += PG) {
      if (j0 > 0) __syncthreads();  // the previous phase's planes are read
      // the 8 planes' tiles straight into LDS (buffer_load ... lds, one dword per lane, agent-
      // coherent): every load of the phase in flight at once, none through registers.  Tile
      // dword d of a plane is double d / 2 (row-major W x W, own points included: the global
      // copy equals the registers here); a wave's 64 lanes fill 64 consecutive dwords.  Points
      // outside an isolated cell get an offset past the buffer's end, which reads 0.
      int zi = z0 + j0;
      asm volatile("" : "+s"(zi));  // offsets formed here, not hoisted out of the CG loop
#pragma unroll 1
      for (int jj = 0; jj < PG; jj++) {
        const long zoff = plane * (zi + jj);
#pragma unroll 1
        for (int d0 = 0; d0 < 2 * W * W; d0 += 256) {
          const int d = d0 + t;
          if (d < 2 * W * W) {
            const int e2 = d >> 1, ly = e2 / W, lx = e2 - ly * W;
            int X = x0 + lx - H, Y = y0 + ly - H;
            bool zero = false;
            if (X < 0 || X >= a.nx) { if (a.px) X = hcgp_wrap(X, a.nx); else zero = true; }
            if (Y < 0 || Y >= a.ny) { if (a.py) Y = hcgp_wrap(Y, a.ny); else zero = true; }
            const unsigned off = zero ? 0xFFFFFFF0u
                                      : (unsigned)((X + (long)a.nx * Y + zoff) * 8 + (d & 1) * 4);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                prs, (__attribute__((address_space(3))) void *)((char *)&tile[jj][0] + 4 * (d0 + (t & ~63))),
                4, off, 0, 0, HCGP_AUX_SC1);
          }
        }
      }
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's tile loads have landed
      __syncthreads();
#pragma unroll
      for (int jj = 0; jj < PG; jj++) {
        const int j = j0 + jj;
        const double *c = &tile[jj][(ty + H) * W + tx + H];
        double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
        for (int m = 1; m <= H; m++) {
          sx += cf[m] * (c[-m] + c[m]);
          sy += cf[H + 1 + m] * (c[-m * W] + c[m * W]);
          const double zm = j - m >= 0 ? e[j - m] : zh[H + j - m][t];
          const double zp = j + m < P ? e[j + m] : zh[H + j + m - P][t];
          sz += cf[2 * (H + 1) + m] * (zm + zp);
        }
        const double sc = a.c0 * e[j] + (sx + sy + sz);
        ap[j] = sc;
        pap += e[j] * sc;
        // finish this plane here: the arithmetic would otherwise sink to the first use of
        // pap, after the last plane, keeping every plane's LDS values live at once
        asm volatile("" : "+v"(pap));
      }
    }
    const double papt = hcgp_allsum(pap, a.slot1, a.bar, G, gen, red4, &ok_sh);
    if (!ok_sh) { status = 2; break; }
    const double alph = r1r / papt;
    double r2 = 0.0;
    const double *bcol = a.b + (long)x + (long)a.nx * y + c0;
#pragma unroll
    for (int j = 0; j < P; j++) {
      vh[j] = vh[j] + alph * e[j];
      av[j] = av[j] + alph * ap[j];
      const double ri = bcol[plane * j] - av[j];
      ap[j] = ri;  // r from here on
      r2 += ri * ri;
    }
    const double r2r = hcgp_allsum(r2, a.slot2, a.bar, G, gen, red4, &ok_sh);
    if (!ok_sh) { status = 2; break; }
    const double dp = sqrt(r2r * a.dvo);
    if (loop >= a.ncgmin && (loop >= a.ncgmax || dp <= a.eps)) break;
    const double beta = (loop % a.ncgres == 0) ? 0.0 : r2r / r1r;
#pragma unroll
    for (int j = 0; j < P; j++) {
      e[j] = ap[j] + beta * e[j];
      hcgp_st(pcol + c0 + plane * j, e[j]);
    }
    r1r = r2r;
    loop++;
  }
  double *vcol = a.vh + (long)x + (long)a.nx * y + c0;
#pragma unroll
  for (int j = 0; j < P; j++) vcol[plane * j] = vh[j];
  if (gi == 0 && t == 0) {
    a.state[0] = status;
    a.state[1] = loop;
  }
}

}  // namespace

// ============================================================================= helpers
Ctx::Ctx(const Grid &g_) : g(g_), s(lib().stream) {}

double *Ctx::buf(const char *tag, size_t count) {
  return (double *)scratch(std::string("ctx.") + tag, count * sizeof(double));
}

int *Ctx::flag(int i) {
  int *f = (int *)scratch("ctx.flag", sizeof(int) * 4);
  return f + i;
}

void Ctx::clear_flag(int i) { HIPCK(hipMemsetAsync(flag(i), 0, sizeof(int), s)); }

int Ctx::read_flag(int i) {
  int h = 0;
  HIPCK(hipMemcpyAsync(&h, flag(i), sizeof(int), hipMemcpyDeviceToHost, s));
  sync();
  return h;
}

void Ctx::proj(const double *psi, long ld, int nb, double *rp, long ld_rp) {
  launch_gather(pl, psi, ld, nb, rp, ld_rp, false, s);
}

void Ctx::hpsi(int ns, const double *veffs, const double *psi, long ld, int nb, const double *rp,
               long ld_rp, double *out, long ld_out, bool accumulate, double *eps_out,
               bool rp_exact, double *pscale_out, double *hscale_out, double *norm_out) {
  const long n1 = (long)pl.nprjmx * pl.num_ppcell;
  double *coef = buf("coef", (size_t)std::max<long>(1, n1) * nb);
  const bool bf = stencil_band_fastest_coef(g, false);
  launch_coef(pl, rp, ld_rp, nb, dij, false, ns, 1.0 / g.dv(), coef, n1, s, bf);
  StencilCall c;
  c.g = &g; c.in = psi; c.ld_in = ld; c.out = out; c.ld_out = ld_out; c.veff = veffs; c.nb = nb;
  c.accumulate = accumulate; c.pl = &pl; c.coef = coef; c.ld_coef = n1; c.coef_band_fastest = bf;
  double *part = nullptr;
  int np = stencil_nparts(g, g.nf, nb);
  const bool pdots = rp_exact && !accumulate && param("nl_proj_dots", 1.0) >= 1.0;
  if (eps_out) {
    // the tiled path appends the expansion's partials (one column from projector space)
    const int np_total = np + (pdots ? 1 : expand_nparts(pl));
    part = buf("hpsi.part", (size_t)np_total * nb);
    launch_fill((long)np_total * nb, 0.0, part, s);
    c.dot_partial = part;
    c.dot_ld = np_total;
    np = np_total;
    if (pdots) { c.rp = rp; c.ld_rp = ld_rp; }
  }
  // psi's column maxima and a bound on out's from the same pass (tiled path)
  const bool tiled = bf && !accumulate;
  const int nx = expand_nparts(pl);
  double *xmax = nullptr;
  unsigned long long *smax = nullptr, *pmax = nullptr;
  const int nq = c.dot_ld ? (int)c.dot_ld : stencil_nparts(g, g.nf, nb);
  double *npart = nullptr;
  if (tiled && norm_out) {  // same layout as the <psi|out> partials
    npart = buf("hpsi.npart", (size_t)nq * nb);
    launch_fill((long)nq * nb, 0.0, npart, s);
    c.norm_partial = npart;
  }
  if (tiled && pscale_out) {
    pmax = (unsigned long long *)buf("hpsi.pmax", (size_t)nb);
    launch_fill(nb, 0.0, (double *)pmax, s);  // +0.0 has all bits zero
    c.in_max = pmax;
  }
  if (tiled && hscale_out) {
    smax = (unsigned long long *)buf("hpsi.smax", (size_t)nb);
    launch_fill(nb, 0.0, (double *)smax, s);  // +0.0 has all bits zero
    c.out_max = smax;
    if (pl.nent > 0) {
      xmax = buf("hpsi.xmax", (size_t)nx * nb);
      c.expand_max = xmax;
    }
  }
  c.tag = "hpsi_stencil";  // the stencil kernel; the projector pass has its own tag
  launch_stencil(c, s);
  if (eps_out) launch_reduce_partials(part, np, nb, g.dv(), eps_out, s);
  if (norm_out) {
    if (npart) launch_reduce_partials(npart, nq, nb, g.dv(), norm_out, s);
    else dots(psi, ld, psi, ld, nb, norm_out);
  }
  if (pscale_out) {
    if (pmax) maxscale(pmax, nb, pscale_out, s);
    else colscale(psi, ld, g.n(), nb, pscale_out, s);
  }
  if (hscale_out) {
    if (smax) launch_bound_scale(nb, smax, xmax, nx, hscale_out, s);
    else colscale(out, ld_out, g.n(), nb, hscale_out, s);
  }
}

bool Ctx::dij_symmetric() {
  if (dij_sym < 0) {
    const long count = (long)pl.nprjmx * pl.nprjmx * pl.nums * pl.natom;
    dij_sym = 1;
    if (dij && count > 0) {
      clear_flag(3);
      hipLaunchKernelGGL(k_dij_asym, dim3(nblk(count, 256)), dim3(256), 0, s, dij, pl.nprjmx,
                         count, flag(3));
      dij_sym = read_flag(3) ? 0 : 1;
    }
  }
  return dij_sym == 1;
}

void Ctx::hpsi_comb(int ns, const double *veffs, const double *x, const double *y,
                    const double *hy, long ld, int nb, const double *ca, const double *cb,
                    const double *rpx, const double *rp, long ld_rp, double *out, double *pap,
                    double *sap, const double *rpy, const double *eh) {
  const long n1 = (long)pl.nprjmx * pl.num_ppcell;
  double *coef = buf("coef", (size_t)std::max<long>(1, n1) * nb);
  // the projector term of ca x alone: cb H y already holds cb's
  launch_coef(pl, rpx, ld_rp, nb, dij, false, ns, 1.0 / g.dv(), coef, n1, s, true);
  const bool pdots = rpy && param("nl_proj_dots", 1.0) >= 1.0;
  // H symmetric (dij symmetric): <y|H p> = <H y|ca x> + cb <y|H y>, so y is not read
  const bool sym = pdots && eh && param("comb_sym", 1.0) >= 1.0 && dij_symmetric();
  const int np = stencil_nparts(g, g.nf, nb, true) + (pdots ? 1 : expand_nparts(pl));
  double *part = buf("hpsi.part", (size_t)2 * np * nb);
  launch_fill((long)2 * np * nb, 0.0, part, s);
  StencilCall c;
  c.g = &g; c.in = x; c.in2 = sym ? nullptr : y; c.ca = ca; c.cb = cb; c.ld_in = ld; c.out = out; c.ld_out = ld;
  c.veff = veffs; c.nb = nb; c.pl = &pl; c.coef = coef; c.ld_coef = n1; c.coef_band_fastest = true;
  c.dot_partial = part; c.dot_partial2 = part + (long)np * nb; c.dot_ld = np;
  if (pdots) { c.rp = sym ? rpx : rp; c.rp2 = sym ? nullptr : rpy; c.ld_rp = ld_rp; }
  c.hin = hy;
  c.tag = "hp_comb_stencil";
  launch_stencil(c, s);
  launch_reduce_partials(part, np, nb, g.dv(), pap, s);
  launch_reduce_partials(part + (long)np * nb, np, nb, g.dv(), sap, s);
  if (sym) hipLaunchKernelGGL(k_comb_sym, dim3(nblk(nb, 64)), dim3(64), 0, s, nb, cb, eh, pap, sap);
}

void Ctx::spsi(const double *psi, long ld, int nb, const double *rp, long ld_rp, double *out,
               long ld_out) {
  if (out != psi) launch_copy2d(g.n(), nb, psi, ld, out, ld_out, s);
  if (q_zero) return;
  const long n1 = (long)pl.nprjmx * pl.num_ppcell;
  double *coef = buf("scoef", (size_t)std::max<long>(1, n1) * nb);
  launch_coef(pl, rp, ld_rp, nb, sss, true, 1, 1.0 / g.dv(), coef, n1, s);
  launch_expand(pl, g.n(), coef, n1, nb, out, ld_out, s);
}

void Ctx::dots(const double *x, long ldx, const double *y, long ldy, int nb, double *res) {
  launch_dots(g.n(), x, ldx, y, ldy, nb, g.dv(), res, s);
}

// rocBLAS dgemm with timing under tag "gemm_f64" (2 m n k flops)
static void dgemm(rocblas_operation ta, rocblas_operation tb, int m, int n, int k, double alpha,
                  const double *A, int lda, const double *B, int ldb, double beta, double *C,
                  int ldc) {
  Timed t("gemm_f64", 2.0 * (double)m * n * k);
  BLASCK(rocblas_dgemm(lib().blas, ta, tb, m, n, k, &alpha, A, lda, B, ldb, &beta, C, ldc));
}

// C(M x M) = alpha A^T B + beta C in FP64 where the consumer reads one triangle only (tri 1:
// on and below the diagonal, 2: on and above): the b (b + 1) / 2 blocks of a b x b partition
// that hold it, b = param "gram_blocks" (default 4) limited to blocks of at least 256 columns
// (narrower blocks ran slower than the full product).  At 2,097,152 x 1024 x 1024: 44 ms
// against 69 ms for the full dgemm.  Entries outside the blocks are left as they are.
static void dgemm_tn_tri(int M, long K, double alpha, const double *A, long lda, const double *B,
                         long ldb, double beta, double *C, long ldc, int tri) {
  const int nbk = tri ? std::max(1, std::min((int)param("gram_blocks", 4.0), M / 256)) : 1;
  if (nbk <= 1) {
    dgemm(rocblas_operation_transpose, rocblas_operation_none, M, M, (int)K, alpha, A, (int)lda, B,
          (int)ldb, beta, C, (int)ldc);
    return;
  }
  for (int bj = 0; bj < nbk; bj++) {
    const int j0 = (int)((long)M * bj / nbk), j1 = (int)((long)M * (bj + 1) / nbk);
    for (int bi = 0; bi < nbk; bi++) {
      if ((tri == 1 && bi < bj) || (tri == 2 && bi > bj)) continue;
      const int i0 = (int)((long)M * bi / nbk), i1 = (int)((long)M * (bi + 1) / nbk);
      dgemm(rocblas_operation_transpose, rocblas_operation_none, i1 - i0, j1 - j0, (int)K, alpha,
            A + lda * i0, (int)lda, B + ldb * j0, (int)ldb, beta, C + i0 + ldc * j0, (int)ldc);
    }
  }
}

// Subspace products: FP64 rocBLAS, or split-FP16 MFMA when param "precision" = 1 and the
// long dimension reaches param "mixed_min_rows" (default 65536: grid-length operands only;
// small products such as <p|psi> rotations stay FP64).
static bool mixed_for(long rows) {
  return param("precision", 0.0) >= 1.0 && rows > 0 &&
         rows >= (long)param("mixed_min_rows", 65536.0);
}
static int mixed_terms() { return param("mixed_terms", 3.0) >= 4.0 ? 4 : 3; }

const double *mixed_scale(const char *tag, const double *X, long ldx, long rows, int cols) {
  if (!mixed_for(rows) || cols <= 0) return nullptr;
  double *sc = (double *)scratch(std::string("scale.") + tag, sizeof(double) * cols);
  colscale(X, ldx, rows, cols, sc, lib().stream);
  return sc;
}

void subspace_tn(int M1, int M2, long K, double alpha, const double *A, long lda,
                 const double *B, long ldb, double beta, double *C, long ldc, const double *sa,
                 const double *sb, const SplitBands *pa, int tri) {
  hipStream_t s = lib().stream;
  if (mixed_for(K)) {
    if (pa) sa = pa->s;
    if (!sa) sa = mixed_scale("tn.a", A, lda, K, M1);
    if (!sb) sb = (B == A && ldb == lda && M2 == M1) ? sa : mixed_scale("tn.b", B, ldb, K, M2);
    gemm_mixed_tn(M1, M2, K, alpha, A, lda, sa, B, ldb, sb, beta, C, ldc, mixed_terms(), s, pa,
                  param("tn_triangle", 1.0) >= 1.0 ? tri : 0);
    return;
  }
  if (tri && M1 == M2 && param("tn_triangle", 1.0) >= 1.0) {
    dgemm_tn_tri(M1, K, alpha, A, lda, B, ldb, beta, C, ldc, tri);
    return;
  }
  dgemm(rocblas_operation_transpose, rocblas_operation_none, M1, M2, (int)K, alpha, A, (int)lda, B,
        (int)ldb, beta, C, (int)ldc);
}

void subspace_nn(long R, int K, int M2, double alpha, const double *X, long ldx, const double *C,
                 long ldc, double beta, double *Y, long ldy, const double *sx,
                 const SplitBands *px, double *colsq, double dv, bool c_upper) {
  hipStream_t s = lib().stream;
  if (mixed_for(R)) {
    if (px) sx = px->s;
    if (!sx) sx = mixed_scale("nn.x", X, ldx, R, K);
    gemm_mixed_nn(R, K, M2, alpha, X, ldx, sx, C, ldc, beta, Y, ldy, mixed_terms(), s, px, colsq, dv,
                  c_upper && param("nn_triangle", 1.0) >= 1.0);
    return;
  }
  if (colsq) throw Error(1, "subspace_nn: column norms come from the split-precision product only");
  dgemm(rocblas_operation_none, rocblas_operation_none, (int)R, M2, K, alpha, X, (int)ldx, C,
        (int)ldc, beta, Y, (int)ldy);
}

static void sub_tn(Ctx &, int M1, int M2, long K, double alpha, const double *A, long lda,
                   const double *B, long ldb, double beta, double *C, long ldc,
                   const double *sa = nullptr, const double *sb = nullptr,
                   const SplitBands *pa = nullptr, int tri = 0) {
  subspace_tn(M1, M2, K, alpha, A, lda, B, ldb, beta, C, ldc, sa, sb, pa, tri);
}
// colsq: dV sum Y^2 per column of the result, from the product's epilogue (split precision)
static void sub_nn(Ctx &c, long R, int K, int M2, double alpha, const double *X, long ldx,
                   const double *C, long ldc, double beta, double *Y, long ldy,
                   const double *sx = nullptr, const SplitBands *px = nullptr,
                   double *colsq = nullptr, bool c_upper = false) {
  subspace_nn(R, K, M2, alpha, X, ldx, C, ldc, beta, Y, ldy, sx, px, colsq, c.g.dv(), c_upper);
}

// Psi split once for the TN and the NN product that follow (split-precision runs only); a
// split scale from the band norms when they are known (|psi| <= sqrt(g / dV)), else from
// the column maxima.
static const SplitBands *split_psi(Ctx &c, const char *tag, const double *psi, int M,
                                   const double *norms, SplitBands &sp) {
  const long n = c.g.n();
  if (!mixed_for(n)) return nullptr;
  const double *sc;
  if (norms) {
    double *b = c.buf((std::string("split.") + tag).c_str(), M);
    normscale(norms, M, c.g.dv(), b, c.s);
    sc = b;
  } else {
    sc = mixed_scale(tag, psi, n, n, M);
  }
  split_bands(psi, n, n, M, sc, sp, c.s);
  return &sp;
}

// Psi split with given column scales (split-precision runs only)
static const SplitBands *split_psi_scaled(Ctx &c, const double *psi, int M, const double *sc,
                                          SplitBands &sp) {
  const long n = c.g.n();
  if (!mixed_for(n)) return nullptr;
  split_bands(psi, n, n, M, sc, sp, c.s);
  return &sp;
}

// After a split-precision rotation the band norms carry ~1e-7 relative errors, which would
// enter the (unnormalised) Rayleigh quotients dV <psi|H|psi> at first order.  Restore
// dV <psi|S|psi> = 1 per band in FP64, scaling S psi, <p|psi> and H psi alike.
// norms: dV <psi|S psi> when already known (from the product that wrote psi).
static void renormalize(Ctx &c, int M, double *psi, double *spsi, double *rsp, double *hpsi,
                        const double *norms = nullptr) {
  const long n = c.g.n(), n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
  double *f = c.buf("renorm.f", M);
  c.clear_flag(2);
  if (norms) HIPCK(hipMemcpyAsync(f, norms, sizeof(double) * M, hipMemcpyDeviceToDevice, c.s));
  else c.dots(psi, n, spsi, n, M, f);
  hipLaunchKernelGGL(k_rsqrt, dim3(nblk(M, 64)), dim3(64), 0, c.s, M, f, c.flag(2));
  if (c.read_flag(2)) throw Error(2, "orthogonalization_r_01 (1): wf has negative norm");
  launch_scale(n, M, f, psi, n, c.s);
  if (spsi != psi) launch_scale(n, M, f, spsi, n, c.s);
  if (hpsi) launch_scale(n, M, f, hpsi, n, c.s);
  if (rsp && n1 > 0) launch_scale(n1, M, f, rsp, n1, c.s);
}

// ============================================================================= Cholesky-QR
// One pass of S-orthonormalisation of bands [0, M): G = dV Psi^T S Psi = R^T R,
// Psi, S Psi, <p|psi> (and H Psi if given) <- X R^-1.  Same Q as Gram-Schmidt in exact
// arithmetic (orthogonalization8f.F90:93-376 with l_inp = 0).
// fp64: the pass runs in FP64 also when the subspace products are split-precision (the
// closing orthonormalisation of the band update: the bands it returns are orthonormal to
// FP64 rounding, not to the split products' ~1e-7).  In FP64, G is one rocBLAS dgemm: at
// 2,097,152 x 1024 dsyrk took 219 ms against about 62 ms for the full dgemm.
void cholqr(Ctx &c, int M, double *psi, double *spsi, double *rsp, double *hpsi,
            const double *norms, bool fp64) {
  const long n = c.g.n(), n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
  rocblas_handle h = lib().blas;
  double *G = c.buf("cholqr.G", (size_t)M * M);
  const double dv = c.g.dv(), one = 1.0, zero = 0.0;
  const bool split = mixed_for(n) && !fp64;
  if (fp64 && mixed_for(n)) {
    // the split-precision rotations before this pass updated <p|psi> and S psi beside psi,
    // each with its own ~1e-7 rounding: recompute both from psi in FP64 so that the pass
    // S-orthonormalises psi itself (and <p|psi>, S psi match the returned bands)
    if (rsp && n1 > 0) c.proj(psi, n, M, rsp, n1);
    if (spsi != psi && rsp) c.spsi(psi, n, M, rsp, n1, spsi, n);
  }
  SplitBands spl;
  const SplitBands *pre = nullptr;
  const double *s_psi = nullptr;
  if (split) {
    // psi split once for G's TN and the first rotation (psi's own, next)
    pre = split_psi(c, "cholqr.psi", psi, M, norms, spl);
    s_psi = pre ? pre->s : nullptr;
    // dpotrf reads the upper triangle only
    sub_tn(c, M, M, n, dv, psi, n, spsi, n, 0.0, G, M, s_psi, nullptr, pre, 2);  // spsi == psi: one split
  } else {
    dgemm_tn_tri(M, n, dv, psi, n, spsi, n, zero, G, M, 2);  // dpotrf reads the upper triangle
  }
  int *info = (int *)scratch("cholqr.info", sizeof(int));
  {
    Timed t("potrf", (double)M * M * M / 3.0);
    BLASCK(rocsolver_dpotrf(h, rocblas_fill_upper, M, G, M, info));
  }
  int hinfo = 0;
  HIPCK(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c.s));
  sync();
  if (hinfo != 0) throw Error(2, "orthogonalization_r_01 (1): wf has negative norm");
  auto tr = [&](double *X, long ldx, long rows) {
    Timed t("trsm_f64", (double)rows * M * M);
    BLASCK(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                         rocblas_diagonal_non_unit, (int)rows, M, &one, G, M, X, (int)ldx));
  };
  if (split) {
    // R^-1 in FP64 (upper triangular), then X <- X R^-1 as a split-FP16 product
    double *Ri = c.buf("cholqr.Ri", (size_t)M * M);
    HIPCK(hipMemcpy2DAsync(Ri, sizeof(double) * M, G, sizeof(double) * M, sizeof(double) * M, M,
                           hipMemcpyDeviceToDevice, c.s));
    {
      Timed t("trtri", (double)M * M * M / 3.0);
      BLASCK(rocsolver_dtrtri(h, rocblas_fill_upper, rocblas_diagonal_non_unit, M, Ri, M, info));
    }
    hipLaunchKernelGGL(k_zero_lower, dim3(nblk((long)M * M, 256)), dim3(256), 0, c.s, M, Ri);
    // in place: the split product reads X only through its pre-split copy
    auto apply = [&](double *X, const double *sx, const SplitBands *px) {
      sub_nn(c, n, M, M, 1.0, X, n, Ri, M, 0.0, X, n, sx, px, nullptr, true);
    };
    // with S psi = psi the product that writes psi also gives its norms
    double *nrm = spsi == psi ? c.buf("cholqr.norms", M) : nullptr;
    sub_nn(c, n, M, M, 1.0, psi, n, Ri, M, 0.0, psi, n, s_psi, pre, nrm, true);  // R^-1 upper
    if (spsi != psi) apply(spsi, nullptr, nullptr);
    if (hpsi) apply(hpsi, nullptr, nullptr);
    if (rsp && n1 > 0) tr(rsp, n1, n1);
    renormalize(c, M, psi, spsi, rsp, hpsi, nrm);
    return;
  }
  tr(psi, n, n);
  if (spsi != psi) tr(spsi, n, n);
  if (rsp && n1 > 0) tr(rsp, n1, n1);
  if (hpsi) tr(hpsi, n, n);
}

// ============================================================================= single-band GS
// orthogonalization8f.F90:316-369: band l (0-based) against bands 0..l-1, then S-normalise.
void gs_one(Ctx &c, int l, double *psi, double *spsi, double *rsp) {
  const long n = c.g.n(), n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
  rocblas_handle h = lib().blas;
  double *av = c.buf("gs1.av", std::max(1, l));
  const double dv = c.g.dv(), zero = 0.0, one = 1.0, mone = -1.0;
  double *v = psi + n * l, *sv = spsi + n * l, *rv = rsp + n1 * l;
  if (l > 0) {
    BLASCK(rocblas_dgemv(h, rocblas_operation_transpose, (int)n, l, &dv, spsi, (int)n, v, 1, &zero,
                         av, 1));
    BLASCK(rocblas_dgemv(h, rocblas_operation_none, (int)n, l, &mone, psi, (int)n, av, 1, &one, v, 1));
    BLASCK(rocblas_dgemv(h, rocblas_operation_none, (int)n, l, &mone, spsi, (int)n, av, 1, &one, sv, 1));
    if (n1 > 0)
      BLASCK(rocblas_dgemv(h, rocblas_operation_none, (int)n1, l, &mone, rsp, (int)n1, av, 1, &one, rv, 1));
  }
  double *vn = c.buf("gs1.vn", 2);
  c.dots(v, n, sv, n, 1, vn);
  hipLaunchKernelGGL(k_rsqrt_norm, dim3(1), dim3(64), 0, c.s, 1, vn, vn + 1, c.flag());
  launch_scale(n, 1, vn + 1, v, n, c.s);
  launch_scale(n, 1, vn + 1, sv, n, c.s);
  if (n1 > 0) launch_scale(n1, 1, vn + 1, rv, n1, c.s);
}

void ortho_r_01(Ctx &c, int M, int l_inp, int nflag, int key_cmpt, double *psi, double *spsi,
                double *rsp) {
  const long n = c.g.n(), n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
  const int l0 = l_inp == 0 ? 0 : l_inp - 1, l1 = l_inp == 0 ? M : l_inp;
  if (nflag == key_cmpt) {
    c.proj(psi + n * l0, n, l1 - l0, rsp + n1 * l0, n1);
    c.spsi(psi + n * l0, n, l1 - l0, rsp + n1 * l0, n1, spsi + n * l0, n);
  }
  if (l_inp == 0) {
    cholqr(c, M, psi, spsi, rsp, nullptr);
  } else {
    c.clear_flag();
    gs_one(c, l_inp - 1, psi, spsi, rsp);
    if (c.read_flag()) throw Error(2, "orthogonalization_r_01 (3): wf has negative norm");
  }
}

// ============================================================================= Rayleigh-Ritz
// rayleigh-ritz8f.F90:22-125: A = dV Psi^T (H Psi), eigh (lower), rotate Psi, S Psi, <p|psi>.
void rayleigh_ritz(Ctx &c, int M, int ns, const double *veffs, bool l_rrb, double *psi,
                   double *spsi, double *rsp, double *hpsi, double *sval, bool keep_h,
                   bool h_scratch) {
  const long n = c.g.n(), n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
  rocblas_handle h = lib().blas;
  // split-precision: the H psi pass also gives psi's column maxima and a bound on those of
  // H psi (their split scales), so neither needs a pass of its own
  double *psc = nullptr, *hsc = nullptr;
  if (l_rrb && mixed_for(n) && param("rr_fused_scales", 1.0) >= 1.0) {
    psc = c.buf("rr.pscale", M);
    hsc = c.buf("rr.hscale", M);
  }
  if (l_rrb) c.hpsi(ns, veffs, psi, n, M, rsp, n1, hpsi, n, false, nullptr, false, psc, hsc);
  double *A = c.buf("rr.A", (size_t)M * M);
  double *E = c.buf("rr.E", (size_t)M);
  const double dv = c.g.dv(), zero = 0.0, one = 1.0;
  // psi split once for this TN and the rotation's NN (no other product runs in between)
  SplitBands spl;
  const SplitBands *pre = psc ? split_psi_scaled(c, psi, M, psc, spl)
                              : split_psi(c, "rr.psi", psi, M, nullptr, spl);
  const double *s_psi = pre ? pre->s : nullptr;
  const double *s_h = hsc ? hsc : mixed_scale("rr.h", hpsi, n, n, M);
  sub_tn(c, M, M, n, dv, psi, n, hpsi, n, zero, A, M, s_psi, s_h, pre, 1);  // dsyevd reads the lower triangle
  int *info = (int *)scratch("rr.info", sizeof(int));
  {
    Timed t("syevd", 4.0 * (double)M * M * M);
    BLASCK(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, M, A, M, sval, E, info));
  }
  int hinfo = 0;
  HIPCK(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c.s));
  sync();
  if (hinfo != 0) throw Error(3, "rayleighritz_diag_r: error in matrix diagonalization");
  launch_canon_sign(M, A, c.s);
  if (h_scratch && !keep_h && mixed_for(n) && !(c.q_zero == false && spsi != psi)) {
    // split products, H psi not needed afterwards: Psi rotated in place (the product reads
    // it only through the pre-split copy), no band-sized temporary or copy
    // (S psi = psi here: the product that writes psi also gives its norms)
    double *nrm = c.buf("rr.norms", M);
    sub_nn(c, n, M, M, one, psi, n, A, M, zero, psi, n, s_psi, pre, nrm);
    if (spsi != psi)
      HIPCK(hipMemcpyAsync(spsi, psi, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
    if (n1 > 0) {
      double *R = c.buf("rr.R", (size_t)n1 * M);
      dgemm(rocblas_operation_none, rocblas_operation_none, (int)n1, M, M, one, rsp, (int)n1, A, M,
            zero, R, (int)n1);
      HIPCK(hipMemcpyAsync(rsp, R, sizeof(double) * n1 * M, hipMemcpyDeviceToDevice, c.s));
    }
    if (param("rr_renorm", 1.0) >= 1.0) renormalize(c, M, psi, spsi, rsp, nullptr, nrm);
    return;
  }
  if (keep_h && mixed_for(n)) {
    // block path, split products: Psi, H Psi (and S Psi) rotated in place (each product reads
    // its X only through the pre-split copy), <p|psi> through a small buffer; psi first, while
    // its shared split is still in place
    sub_nn(c, n, M, M, one, psi, n, A, M, zero, psi, n, s_psi, pre);
    sub_nn(c, n, M, M, one, hpsi, n, A, M, zero, hpsi, n, s_h);
    if (c.q_zero && spsi != psi)
      HIPCK(hipMemcpyAsync(spsi, psi, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
    else if (spsi != psi)
      sub_nn(c, n, M, M, one, spsi, n, A, M, zero, spsi, n);
    if (n1 > 0) {
      double *R = c.buf("rr.R", (size_t)n1 * M);
      dgemm(rocblas_operation_none, rocblas_operation_none, (int)n1, M, M, one, rsp, (int)n1, A, M,
            zero, R, (int)n1);
      HIPCK(hipMemcpyAsync(rsp, R, sizeof(double) * n1 * M, hipMemcpyDeviceToDevice, c.s));
    }
    renormalize(c, M, psi, spsi, rsp, hpsi);
    return;
  }
  // rotations through hsvre as the reference does (:96-119)
  double *W = hpsi;
  double *Wh = nullptr;
  if (keep_h) {  // block path: rotate H Psi too, through a separate buffer
    Wh = c.buf("tmp.nM", (size_t)n * M);  // shared with Cholesky-QR (never live together)
    sub_nn(c, n, M, M, one, hpsi, n, A, M, zero, Wh, n, s_h);
  }
  sub_nn(c, n, M, M, one, psi, n, A, M, zero, W, n, s_psi, keep_h ? nullptr : pre);
  HIPCK(hipMemcpyAsync(psi, W, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
  if (c.q_zero && spsi != psi) {
    HIPCK(hipMemcpyAsync(spsi, psi, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
  } else if (spsi != psi) {
    sub_nn(c, n, M, M, one, spsi, n, A, M, zero, W, n);
    HIPCK(hipMemcpyAsync(spsi, W, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
  }
  if (n1 > 0) {
    double *R = c.buf("rr.R", (size_t)n1 * M);
    dgemm(rocblas_operation_none, rocblas_operation_none, (int)n1, M, M, one, rsp, (int)n1, A, M,
          zero, R, (int)n1);
    HIPCK(hipMemcpyAsync(rsp, R, sizeof(double) * n1 * M, hipMemcpyDeviceToDevice, c.s));
  }
  if (keep_h) {
    HIPCK(hipMemcpyAsync(hpsi, Wh, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
  } else if (!(c.q_zero && spsi != psi)) {
    // W already holds S Psi rotated
  } else {
    HIPCK(hipMemcpyAsync(hpsi, spsi, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
  }
  if (mixed_for(n)) renormalize(c, M, psi, spsi, rsp, keep_h ? hpsi : nullptr);
}

// ============================================================================= kscg_r
// Band-sequential preconditioned CG exactly as kscg8f.F90:98-472.
void kscg_sequential(Ctx &c, const KscgArgs &a) {
  const long n = c.g.n(), n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
  const int M = a.neigmx;
  rocblas_handle h = lib().blas;
  double *rvecre = c.buf("cg.rvecre", n), *pvecre = c.buf("cg.pvecre", n), *pvre = c.buf("cg.pvre", n);
  double *rvre = c.buf("cg.rvre", n), *spvre = c.buf("cg.spvre", n), *avre = c.buf("cg.avre", n);
  double *asvre = c.buf("cg.asvre", n);
  double *rrpsep = c.buf("cg.rrpsep", std::max<long>(1, n1)), *rppsep = c.buf("cg.rppsep", std::max<long>(1, n1));
  double *rpsep = c.buf("cg.rpsep", std::max<long>(1, n1));
  double *avreal = c.buf("cg.avreal", M);
  // device scalars: 0 r1rr, 1 r2rr, 2 pap, 3 sapre, 4 vecnor, 5 tmp, 6 alph, 7 beta, 8 t, 9 eps
  double *sc = c.buf("cg.sc", 16);
  const double dv = c.g.dv(), zero = 0.0, one = 1.0, mone = -1.0;
  const int key = a.northo == 2 ? a.key_cmpt : a.key_nocmpt;
  std::vector<double> res_h(1);
  c.clear_flag(1);
  for (int l = 0; l < M; l++) {
    double *psi = a.psi + n * l, *spsi = a.spsi + n * l, *hp = a.hpsi + n * l, *rps = a.rsp + n1 * l;
    for (int loop = 0; loop < a.northo; loop++) ortho_r_01(c, M, l + 1, key, a.key_cmpt, a.psi, a.spsi, a.rsp);
    if (n1 > 0) HIPCK(hipMemcpyAsync(rpsep, rps, sizeof(double) * n1, hipMemcpyDeviceToDevice, c.s));
    c.hpsi(a.ns, a.veffs, psi, n, 1, rpsep, n1, hp, n, false, nullptr);
    int iiter = 0;
    bool conv = false;
    while (iiter < a.nsdmax && !conv) {
      iiter++;
      bool doreset = (iiter == 1) || (iiter == 10) || ((iiter - 1) % 200 == 0);
      if (doreset) {
        launch_fill(1, 1.0, sc + 1, c.s);
        launch_fill(n, 0.0, pvecre, c.s);
        launch_fill(n, 0.0, rvecre, c.s);
        launch_fill(n, 0.0, pvre, c.s);
        launch_fill(n, 0.0, spvre, c.s);
        if (n1 > 0) launch_fill(n1, 0.0, rppsep, c.s);
      }
      HIPCK(hipMemcpyAsync(sc + 0, sc + 1, sizeof(double), hipMemcpyDeviceToDevice, c.s));
      double *ev = a.sval + l;
      c.dots(psi, n, hp, n, 1, ev);
      // r = -(h - eps s)
      launch_axpby(n, 1, nullptr, -1.0, hp, n, ev, 1.0, spsi, n, rvecre, n, c.s);
      if (a.nprecon_cg == 1) launch_precond(c.g, rvecre, n, false, 1, rvre, n, 1, c.s);
      else HIPCK(hipMemcpyAsync(rvre, rvecre, sizeof(double) * n, hipMemcpyDeviceToDevice, c.s));
      for (int loop = 0; loop < a.northo; loop++) {
        BLASCK(rocblas_dgemv(h, rocblas_operation_transpose, (int)n, l + 1, &dv, a.spsi, (int)n, rvre,
                             1, &zero, avreal, 1));
        BLASCK(rocblas_dgemv(h, rocblas_operation_none, (int)n, l + 1, &mone, a.psi, (int)n, avreal, 1,
                             &one, rvre, 1));
      }
      c.proj(rvre, n, 1, rpsep, n1);
      if (n1 > 0) HIPCK(hipMemcpyAsync(rrpsep, rpsep, sizeof(double) * n1, hipMemcpyDeviceToDevice, c.s));
      c.spsi(rvre, n, 1, rrpsep, n1, asvre, n);
      c.dots(rvecre, n, rvre, n, 1, sc + 1);
      HIPCK(hipMemcpyAsync(a.resid + l, sc + 1, sizeof(double), hipMemcpyDeviceToDevice, c.s));
      if (a.nkscg == 1) {
        hipLaunchKernelGGL(k_ratio, dim3(1), dim3(64), 0, c.s, 1, sc + 1, sc + 0, 0, sc + 8);
        launch_axpby(n, 1, nullptr, 1.0, rvre, n, sc + 8, 1.0, pvre, n, pvre, n, c.s);
        HIPCK(hipMemcpyAsync(pvecre, pvre, sizeof(double) * n, hipMemcpyDeviceToDevice, c.s));
        launch_axpby(n, 1, nullptr, 1.0, asvre, n, sc + 8, 1.0, spvre, n, spvre, n, c.s);
        if (n1 > 0) launch_axpby(n1, 1, nullptr, 1.0, rrpsep, n1, sc + 8, 1.0, rppsep, n1, rppsep, n1, c.s);
      } else {
        HIPCK(hipMemcpyAsync(pvecre, rvre, sizeof(double) * n, hipMemcpyDeviceToDevice, c.s));
        HIPCK(hipMemcpyAsync(spvre, asvre, sizeof(double) * n, hipMemcpyDeviceToDevice, c.s));
        if (n1 > 0) HIPCK(hipMemcpyAsync(rppsep, rrpsep, sizeof(double) * n1, hipMemcpyDeviceToDevice, c.s));
      }
      if (n1 > 0) HIPCK(hipMemcpyAsync(rpsep, rppsep, sizeof(double) * n1, hipMemcpyDeviceToDevice, c.s));
      HIPCK(hipMemcpyAsync(asvre, spvre, sizeof(double) * n, hipMemcpyDeviceToDevice, c.s));
      for (int loop = 0; loop < a.northo; loop++) {
        BLASCK(rocblas_dgemv(h, rocblas_operation_transpose, (int)n, l + 1, &dv, a.spsi, (int)n,
                             pvecre, 1, &zero, avreal, 1));
        BLASCK(rocblas_dgemv(h, rocblas_operation_none, (int)n, l + 1, &mone, a.psi, (int)n, avreal, 1,
                             &one, pvecre, 1));
        BLASCK(rocblas_dgemv(h, rocblas_operation_none, (int)n, l + 1, &mone, a.spsi, (int)n, avreal, 1,
                             &one, asvre, 1));
        if (n1 > 0)
          BLASCK(rocblas_dgemv(h, rocblas_operation_none, (int)n1, l + 1, &mone, a.rsp, (int)n1, avreal,
                               1, &one, rpsep, 1));
      }
      c.dots(pvecre, n, asvre, n, 1, sc + 4);
      hipLaunchKernelGGL(k_rsqrt_norm, dim3(1), dim3(64), 0, c.s, 1, sc + 4, sc + 5, c.flag(1));
      launch_scale(n, 1, sc + 5, pvecre, n, c.s);
      launch_scale(n, 1, sc + 5, asvre, n, c.s);
      if (n1 > 0) launch_scale(n1, 1, sc + 5, rpsep, n1, c.s);
      c.hpsi(a.ns, a.veffs, pvecre, n, 1, rpsep, n1, avre, n, false, sc + 2);
      c.dots(psi, n, avre, n, 1, sc + 3);
      hipLaunchKernelGGL(k_kscg05, dim3(1), dim3(64), 0, c.s, 1, ev, sc + 2, sc + 3, nullptr, sc + 6,
                         sc + 7);
      launch_axpby(n, 1, sc + 6, 1.0, psi, n, sc + 7, 1.0, pvecre, n, psi, n, c.s);
      launch_axpby(n, 1, sc + 6, 1.0, hp, n, sc + 7, 1.0, avre, n, hp, n, c.s);
      launch_axpby(n, 1, sc + 6, 1.0, spsi, n, sc + 7, 1.0, asvre, n, spsi, n, c.s);
      if (n1 > 0) launch_axpby(n1, 1, sc + 6, 1.0, rps, n1, sc + 7, 1.0, rpsep, n1, rps, n1, c.s);
      HIPCK(hipMemcpyAsync(res_h.data(), a.resid + l, sizeof(double), hipMemcpyDeviceToHost, c.s));
      sync();
      conv = std::sqrt(res_h[0]) < a.epssd;
      *a.ksconv = *a.ksconv && (conv || iiter < a.nsdmax);
    }
    *a.ksitmax = std::max(*a.ksitmax, iiter);
  }
  if (c.read_flag(1)) throw Error(2, "kscg_r: pvec has negative norm");
}

// Block CG: every band advances at once; the reference's per-band projections against
// psi_1..l become one projection against the whole block (two GEMMs), the 2x2 Rayleigh-
// Ritz step of kscg_r_05 runs per band, and Cholesky-QR restores orthonormality at the end.
void kscg_block(Ctx &c, const KscgArgs &a, bool ortho_first, bool keep_h) {
  const long n = c.g.n(), n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
  const int M = a.neigmx;
  const double dv = c.g.dv(), zero = 0.0, one = 1.0, mone = -1.0;
  if (ortho_first)
    for (int loop = 0; loop < a.northo; loop++) cholqr(c, M, a.psi, a.spsi, a.rsp, nullptr);
  double *KR = c.buf("bcg.KR", (size_t)n * M);     // K r, then p
  double *PA = c.buf("bcg.PA", (size_t)n * M);     // accumulated direction
  double *HP = c.buf("bcg.HP", (size_t)n * M);     // H p
  double *SK = c.q_zero ? nullptr : c.buf("bcg.SK", (size_t)n * M);
  double *rpp = c.buf("bcg.rpp", (size_t)std::max<long>(1, n1) * M);
  double *rpx = c.buf("bcg.rpx", (size_t)std::max<long>(1, n1) * M);
  double *C = c.buf("bcg.C", (size_t)M * M);
  double *sc = c.buf("bcg.sc", (size_t)M * 12);
  double *rnew = sc, *rold = sc + M, *tb = sc + 2 * M, *dd = sc + 3 * M, *pp = sc + 4 * M;
  double *gg = sc + 5 * M, *fac = sc + 6 * M, *mdf = sc + 7 * M, *pap = sc + 8 * M;
  double *sap = sc + 9 * M, *al = sc + 10 * M, *be = sc + 11 * M;
  // S = 1: fuse the direction update with its dot products and the band update with the next
  // iteration's <psi|psi> and <psi|H psi> (same arithmetic, fewer passes over HBM)
  const bool fused = c.q_zero;
  // the combination stencil (H p = fac H P + mdf H psi, p never stored) needs the tiled path
  const bool comb = fused && stencil_band_fastest_coef(c.g, false) && param("bcg_comb", 1.0) >= 1.0;
  double *eh_next = c.buf("bcg.eh", (size_t)M);
  int *active = (int *)scratch("bcg.active", sizeof(int) * M);
  std::vector<int> act_h(M, 1);
  HIPCK(hipMemcpyAsync(active, act_h.data(), sizeof(int) * M, hipMemcpyHostToDevice, c.s));
  std::vector<double> res_h(M);
  c.clear_flag(1);
  // H Psi for all bands with their <p|psi>
  // H Psi; with S = 1 the same pass gives the first iteration's dV <psi|H psi> and
  // dV <psi|psi> (no passes of their own)
  double *eh0 = nullptr, *gg0 = nullptr;
  if (fused) {
    eh0 = c.buf("bcg.eh0", (size_t)M);
    gg0 = c.buf("bcg.gg0", (size_t)M);
  }
  c.hpsi(a.ns, a.veffs, a.psi, n, M, a.rsp, n1, a.hpsi, n, false, eh0, false, nullptr, nullptr, gg0);
  // With H symmetric the band update of one iteration and the residual preconditioner of the
  // next run as one stencil pass (launch_update_precond), with the next eps formed as
  // al^2 eps + 2 al be <psi|H p> + be^2 <p|H p>.  That pass writes the new bands to the other
  // buffer of a pair, so psi and H psi alternate between the caller's arrays and bcg.psi2 /
  // bcg.h2; the last band update writes back into the caller's.
  const bool fuse_upd = comb && a.spsi == a.psi && a.nsdmax > 1 &&
                        param("bcg_fuse_update", 1.0) >= 1.0 && c.dij_symmetric();
  double *psi = a.psi, *hpsi = a.hpsi, *spsi = a.spsi;
  double *psi2 = fuse_upd ? c.buf("bcg.psi2", (size_t)n * M) : nullptr;
  double *h2 = fuse_upd ? c.buf("bcg.h2", (size_t)n * M) : nullptr;
  double *rnext = c.buf("bcg.rnext", (size_t)M), *epsn = c.buf("bcg.epsn", (size_t)M);
  bool kr_ready = false;  // K r of the current bands (and rnext, krmax) came with the update
  int iiter = 0;
  bool all_conv = false;
  while (iiter < a.nsdmax && !all_conv) {
    iiter++;
    const bool doreset = (iiter == 1) || (iiter == 10) || ((iiter - 1) % 200 == 0);
    HIPCK(hipMemcpyAsync(rold, rnew, sizeof(double) * M, hipMemcpyDeviceToDevice, c.s));
    // dV <psi|H psi> when S = 1: from the previous iteration's fused band update (the first
    // iteration's from the H psi pass)
    if (fused) {
      HIPCK(hipMemcpyAsync(a.sval, iiter > 1 ? eh_next : eh0, sizeof(double) * M,
                           hipMemcpyDeviceToDevice, c.s));
    } else {
      c.dots(psi, n, hpsi, n, M, a.sval);
    }
    if (fused && iiter == 1)  // later: from the band update
      HIPCK(hipMemcpyAsync(gg, gg0, sizeof(double) * M, hipMemcpyDeviceToDevice, c.s));
    // K r with r = -(H psi - eps S psi), and <r|Kr> (unprojected): residual measure and CG ratio;
    // for split-precision products also the per-band max |K r|
    const bool mixed = mixed_for(n);
    unsigned long long *krmax = mixed ? (unsigned long long *)c.buf("bcg.krmax", M) : nullptr;
    if (kr_ready) {
      HIPCK(hipMemcpyAsync(rnew, rnext, sizeof(double) * M, hipMemcpyDeviceToDevice, c.s));
    } else {
      const int np = stencil_nparts(c.g, 1, M);
      double *part = c.buf("bcg.kpart", (size_t)np * M);
      launch_precond_resid_dot(c.g, hpsi, spsi, n, a.sval, -1.0, KR, M, part, c.s, krmax);
      launch_reduce_partials(part, np, M, c.g.dv(), rnew, c.s);
    }
    // Kr orthogonal to the block in the S metric: C = dV (S Psi)^T Kr ; Kr -= Psi C.  Split
    // scales without passes over the bands: Psi's from dV <psi|psi> (S = 1), K r's from the
    // fused maxima.
    const double *s_kr = nullptr;
    SplitBands spl;
    const SplitBands *pre = nullptr;
    if (mixed) {
      double *sk = c.buf("bcg.skr", M);
      maxscale(krmax, M, sk, c.s);
      s_kr = sk;
      // S = 1: S psi equals psi, so one split of psi serves both products
      if (c.q_zero) pre = split_psi(c, "bcg.psi", psi, M, fused ? gg : nullptr, spl);
    }
    // bcg_proj 1 (default): band b's K r against bands 1..b, the set kscg_r projects on
    // (kscg8f.F90:246-261); C is then upper triangular, so the TN product computes only the
    // tiles on and above the diagonal and the rotation skips the zero stages.  0: all bands.
    const bool lower_only = param("bcg_proj", 1.0) >= 1.0;
    const int tri = lower_only ? 2 : 0;
    auto zero_lower = [&]() {
      if (lower_only)
        hipLaunchKernelGGL(k_zero_lower, dim3(nblk((long)M * M, 256)), dim3(256), 0, c.s, M, C);
    };
    if (pre) {
      sub_tn(c, M, M, n, dv, psi, n, KR, n, zero, C, M, nullptr, s_kr, pre, tri);
      zero_lower();
      sub_nn(c, n, M, M, mone, psi, n, C, M, one, KR, n, nullptr, pre, nullptr, lower_only);
    } else {
      const double *s_psi = mixed_scale("bcg.psi", psi, n, n, M);
      sub_tn(c, M, M, n, dv, spsi, n, KR, n, zero, C, M, spsi == psi ? s_psi : nullptr, s_kr,
             nullptr, tri);
      zero_lower();
      sub_nn(c, n, M, M, mone, psi, n, C, M, one, KR, n, s_psi, nullptr, nullptr, lower_only);
    }
    // P = Kr + beta P
    hipLaunchKernelGGL(k_ratio, dim3(nblk(M, 64)), dim3(64), 0, c.s, M, rnew, rold, doreset ? 1 : 0, tb);
    if (fused) {
      // P = Kr + beta P together with d = dV <S psi|P> and dV <P|P> (S = 1) in one pass
      launch_pupdate_dots(n, M, KR, doreset ? nullptr : tb, PA, n, spsi, dv, dd, pp, c.s);
    } else {
      if (doreset) launch_copy2d(n, M, KR, n, PA, n, c.s);
      else launch_axpby(n, M, nullptr, 1.0, KR, n, tb, 1.0, PA, n, PA, n, c.s);
      // p = (P - d psi)/|.|, d = dV <S psi|P>
      const double *SPA = PA;
      if (!c.q_zero) {
        c.proj(PA, n, M, rpp, n1);
        c.spsi(PA, n, M, rpp, n1, SK, n);
        SPA = SK;
      }
      c.dots(spsi, n, PA, n, M, dd);
      c.dots(PA, n, SPA, n, M, pp);
      c.dots(psi, n, spsi, n, M, gg);
    }
    hipLaunchKernelGGL(k_pnorm, dim3(nblk(M, 64)), dim3(64), 0, c.s, M, dd, pp, gg, fac, mdf, c.flag(1));
    if (comb) {
      // p = fac P + mdf psi is never stored: <proj|p> = fac <proj|P> + mdf <proj|psi> (rsp),
      // and one stencil pass over P gives H p = fac H P + mdf H psi (H psi is maintained),
      // dV <p|H p> and dV <psi|H p>; the projector coefficients are those of fac P (rpx)
      c.proj(PA, n, M, rpp, n1);
      if (n1 > 0) {
        launch_axpby(n1, M, fac, 1.0, rpp, n1, nullptr, 0.0, rpp, n1, rpx, n1, c.s);
        launch_axpby(n1, M, nullptr, 1.0, rpx, n1, mdf, 1.0, a.rsp, n1, rpp, n1, c.s);
      }
      c.hpsi_comb(a.ns, a.veffs, PA, psi, hpsi, n, M, fac, mdf, rpx, rpp, n1, HP, pap, sap,
                  a.rsp, a.sval);
    } else {
      launch_axpby(n, M, fac, 1.0, PA, n, mdf, 1.0, psi, n, KR, n, c.s);  // KR <- p
      c.proj(KR, n, M, rpp, n1);
      c.hpsi(a.ns, a.veffs, KR, n, M, rpp, n1, HP, n, false, pap, true);
      c.dots(psi, n, HP, n, M, sap);
    }
    hipLaunchKernelGGL(k_kscg05, dim3(nblk(M, 64)), dim3(64), 0, c.s, M, a.sval, pap, sap, active, al, be);
    if (comb) {
      // psi <- al psi + be p = (al + be mdf) psi + (be fac) P; H psi <- al H psi + be H p;
      // with dV <psi|psi> and dV <psi|H psi> of the new bands
      double *a2 = c.buf("bcg.a2", (size_t)2 * M), *c2 = a2 + M;
      hipLaunchKernelGGL(k_comb_coef, dim3(nblk(M, 64)), dim3(64), 0, c.s, M, al, be, fac, mdf, a2, c2);
      if (fuse_upd && iiter < a.nsdmax) {
        // this update and the next iteration's K r in one pass, into the other buffers
        hipLaunchKernelGGL(k_eps_next, dim3(nblk(M, 64)), dim3(64), 0, c.s, M, al, be, a.sval, sap,
                           pap, epsn);
        double *po = psi == a.psi ? psi2 : a.psi, *ho = hpsi == a.hpsi ? h2 : a.hpsi;
        const int np = update_precond_nparts(c.g, M);
        double *part = c.buf("bcg.upart", (size_t)3 * np * M);
        launch_update_precond(c.g, psi, PA, hpsi, HP, n, a2, c2, al, be, epsn, -1.0, po, ho, KR, M,
                              part, part + (long)np * M, part + 2L * np * M, np, krmax, c.s);
        launch_reduce_partials(part, np, M, dv, rnext, c.s);
        launch_reduce_partials(part + (long)np * M, np, M, dv, gg, c.s);
        launch_reduce_partials(part + 2L * np * M, np, M, dv, eh_next, c.s);
        psi = po;
        hpsi = ho;
        spsi = psi;
        kr_ready = true;
      } else {
        // in place, or back into the caller's arrays after fused updates
        launch_pair_update_dots(n, M, a2, c2, psi, PA, hpsi, HP, n, dv, gg, eh_next, c.s, al, be,
                                a.psi, a.hpsi);
        psi = a.psi;
        hpsi = a.hpsi;
        spsi = a.spsi;
        kr_ready = false;
      }
    } else if (fused) {
      // psi, H psi updated together, with dV <psi|psi> and dV <psi|H psi> of the new bands
      launch_pair_update_dots(n, M, al, be, psi, KR, hpsi, HP, n, dv, gg, eh_next, c.s);
    } else {
      launch_axpby(n, M, al, 1.0, psi, n, be, 1.0, KR, n, psi, n, c.s);
      launch_axpby(n, M, al, 1.0, hpsi, n, be, 1.0, HP, n, hpsi, n, c.s);
    }
    if (a.spsi != a.psi) {
      if (c.q_zero) {
        HIPCK(hipMemcpyAsync(a.spsi, a.psi, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
      } else {
        c.spsi(KR, n, M, rpp, n1, SK, n);
        launch_axpby(n, M, al, 1.0, a.spsi, n, be, 1.0, SK, n, a.spsi, n, c.s);
      }
    }
    if (n1 > 0) launch_axpby(n1, M, al, 1.0, a.rsp, n1, be, 1.0, rpp, n1, a.rsp, n1, c.s);
    HIPCK(hipMemcpyAsync(a.resid, rnew, sizeof(double) * M, hipMemcpyDeviceToDevice, c.s));
    HIPCK(hipMemcpyAsync(res_h.data(), rnew, sizeof(double) * M, hipMemcpyDeviceToHost, c.s));
    sync();
    all_conv = true;
    for (int l = 0; l < M; l++) {
      const bool cv = std::sqrt(res_h[l]) < a.epssd;
      if (cv) act_h[l] = 0;
      all_conv = all_conv && cv;
      *a.ksconv = *a.ksconv && (cv || iiter < a.nsdmax);
    }
    HIPCK(hipMemcpyAsync(active, act_h.data(), sizeof(int) * M, hipMemcpyHostToDevice, c.s));
  }
  if (psi != a.psi) {  // ended after a fused update: the bands back into the caller's arrays
    HIPCK(hipMemcpyAsync(a.psi, psi, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
    HIPCK(hipMemcpyAsync(a.hpsi, hpsi, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
  }
  *a.ksitmax = std::max(*a.ksitmax, iiter);
  if (c.read_flag(1)) throw Error(2, "kscg_r: pvec has negative norm");
  Timed ph("phase_gs", 0.0);
  // the last pass in FP64 (param final_ortho_fp64, default 1): the returned bands are
  // orthonormal to FP64 rounding also when the subspace products are split-precision
  const bool fp64_last = param("final_ortho_fp64", 1.0) >= 1.0;
  for (int loop = 0; loop < a.northo; loop++)  // gg: dV <psi|psi> from the last band update
    cholqr(c, M, a.psi, a.spsi, a.rsp, keep_h ? a.hpsi : nullptr,
           fused && loop == 0 && iiter > 0 ? gg : nullptr, fp64_last && loop == a.northo - 1);
}

// ============================================================================= ksdiis_r
// ksdiis8f.F90:8-712 with nitermax = 2.  Bands are independent inside a pass, so all bands
// run together; each band's arithmetic is the reference's, expressed on (phi0, K r).
void ksdiis(Ctx &c, const DiisArgs &a) {
  const long n = c.g.n(), n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
  const int M = a.neigmx;
  double *KR = c.buf("diis.KR", (size_t)n * M);
  double *HK = c.buf("diis.HK", (size_t)n * M);
  double *SK = c.q_zero ? KR : c.buf("diis.SK", (size_t)n * M);
  double *rpk = c.buf("diis.rpk", (size_t)std::max<long>(1, n1) * M);
  double *sc = c.buf("diis.sc", (size_t)M * 16);
  double *lam = sc, *ca = sc + M, *cb = sc + 2 * M, *ev = sc + 3 * M, *dots = sc + 4 * M;  // dots: 10M
  const int np = (int)std::min<long>(1024, std::max<long>(1, (n + 8191) / 8192));
  double *part = c.buf("diis.part", (size_t)np * M * 10);
  std::vector<double> res_h(M);
  for (int jiter = 1; jiter < a.ndiismax; jiter++) {  // ksdiis8f.F90:95-97
    // H phi0 and eps
    c.hpsi(a.ns, a.veffs, a.psi, n, M, a.rsp, n1, a.hpsi, n, false, ev);
    // K r0, r0 = H phi0 - eps S phi0
    if (a.nprecon == 1)
      launch_precond_resid(c.g, a.hpsi, a.spsi, n, ev, 1.0, nullptr, KR, M, c.s);
    else
      launch_axpby(n, M, nullptr, 1.0, a.hpsi, n, ev, -1.0, a.spsi, n, KR, n, c.s);
    c.proj(KR, n, M, rpk, n1);
    c.hpsi(a.ns, a.veffs, KR, n, M, rpk, n1, HK, n, false, nullptr);
    if (!c.q_zero) c.spsi(KR, n, M, rpk, n1, SK, n);
    hipLaunchKernelGGL(k_diis_dots, dim3(np, M), dim3(256), 0, c.s, n, a.psi, a.hpsi, a.spsi, KR,
                       HK, SK, n, ev, part, np);
    hipLaunchKernelGGL(k_diis_reduce, dim3(M * 10), dim3(256), 0, c.s, part, np, dots);
    hipLaunchKernelGGL(k_diis_step, dim3(nblk(M, 64)), dim3(64), 0, c.s, M, dots, c.g.dv(),
                       a.alambda_diis, a.alambda_min, a.alambda_max, a.eps_eig_diis, a.sval, lam,
                       ca, cb, a.resid);
    launch_axpby(n, M, ca, 1.0, a.psi, n, cb, 1.0, KR, n, a.psi, n, c.s);
    launch_axpby(n, M, ca, 1.0, a.hpsi, n, cb, 1.0, HK, n, a.hpsi, n, c.s);
    if (a.spsi != a.psi) {
      if (c.q_zero)
        HIPCK(hipMemcpyAsync(a.spsi, a.psi, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
      else
        launch_axpby(n, M, ca, 1.0, a.spsi, n, cb, 1.0, SK, n, a.spsi, n, c.s);
    }
    if (n1 > 0) launch_axpby(n1, M, ca, 1.0, a.rsp, n1, cb, 1.0, rpk, n1, a.rsp, n1, c.s);
    HIPCK(hipMemcpyAsync(res_h.data(), a.resid, sizeof(double) * M, hipMemcpyDeviceToHost, c.s));
    sync();
    for (int l = 0; l < M; l++) *a.ksconv = *a.ksconv && (std::sqrt(res_h[l]) <= a.epssd);
    *a.ksitmax = std::max(*a.ksitmax, 1);
    if (a.l_hsv) {
      for (int i = 0; i < a.northo; i++) ortho_r_01(c, M, 0, a.key_cmpt, a.key_cmpt, a.psi, a.spsi, a.rsp);
      c.hpsi(a.ns, a.veffs, a.psi, n, M, a.rsp, n1, a.hpsi, n, false, nullptr);
    }
  }
}

// ============================================================================= reorderstates
void reorder(Ctx &c, int M, double *sval, double *resid, double *psi, double *spsi, double *rsp,
             double *hpsi, bool move_h, bool h_scratch) {
  const long n = c.g.n(), n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
  std::vector<double> sv(M), rs(M);
  HIPCK(hipMemcpyAsync(sv.data(), sval, sizeof(double) * M, hipMemcpyDeviceToHost, c.s));
  HIPCK(hipMemcpyAsync(rs.data(), resid, sizeof(double) * M, hipMemcpyDeviceToHost, c.s));
  sync();
  std::vector<int> lank(M);
  lank[0] = 1;
  for (int j = 1; j < M; j++) {
    lank[j] = 1;
    for (int i = 0; i < j; i++) {
      if (sv[i] > sv[j]) lank[i]++;
      else lank[j]++;
    }
  }
  std::vector<int> perm(M);
  std::vector<double> st(M), rt(M);
  for (int j = 0; j < M; j++) {
    perm[j] = lank[j] - 1;
    st[lank[j] - 1] = sv[j];
    rt[lank[j] - 1] = rs[j];
  }
  int *pd = (int *)scratch("reo.perm", sizeof(int) * M);
  HIPCK(hipMemcpyAsync(pd, perm.data(), sizeof(int) * M, hipMemcpyHostToDevice, c.s));
  HIPCK(hipMemcpyAsync(sval, st.data(), sizeof(double) * M, hipMemcpyHostToDevice, c.s));
  HIPCK(hipMemcpyAsync(resid, rt.data(), sizeof(double) * M, hipMemcpyHostToDevice, c.s));
  if (n1 > 0) {
    double *R = c.buf("reo.R", (size_t)n1 * M);
    launch_permute_cols(n1, M, pd, rsp, n1, R, n1, c.s);
    HIPCK(hipMemcpyAsync(rsp, R, sizeof(double) * n1 * M, hipMemcpyDeviceToDevice, c.s));
  }
  if (!move_h && !h_scratch) {  // hsvre is the work array (reference): ends as the permuted S psi
    launch_permute_cols(n, M, pd, psi, n, hpsi, n, c.s);
    HIPCK(hipMemcpyAsync(psi, hpsi, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
    if (spsi != psi) {  // S psi may alias psi (S = 1); then hpsi already holds it permuted
      launch_permute_cols(n, M, pd, spsi, n, hpsi, n, c.s);
      HIPCK(hipMemcpyAsync(spsi, hpsi, sizeof(double) * n * M, hipMemcpyDeviceToDevice, c.s));
    }
    sync();
    return;
  }
  // block path: only the bands that change place move (H psi with them when it is kept)
  std::vector<int> mv;
  for (int j = 0; j < M; j++)
    if (perm[j] != j) mv.push_back(j);
  if (!mv.empty()) {
    const int cnt = (int)mv.size();
    std::vector<int> lists(2 * cnt);
    for (int q = 0; q < cnt; q++) {
      lists[q] = mv[q];
      lists[cnt + q] = perm[mv[q]];
    }
    int *ld = (int *)scratch("reo.lists", sizeof(int) * 2 * cnt);
    HIPCK(hipMemcpyAsync(ld, lists.data(), sizeof(int) * 2 * cnt, hipMemcpyHostToDevice, c.s));
    double *tmp = c.buf("reo.tmp", (size_t)n * cnt);
    auto move = [&](double *x) {
      launch_copy_cols(n, cnt, ld, nullptr, x, n, tmp, n, c.s);
      launch_copy_cols(n, cnt, nullptr, ld + cnt, tmp, n, x, n, c.s);
    };
    move(psi);
    if (spsi != psi) move(spsi);  // S psi may alias psi (S = 1); move it only once
    if (move_h) move(hpsi);
  }
  sync();
}

// ============================================================================= Hartree
// scf_hartree8f.F90:11-812: b = 2 pi rho (+ Dirichlet planes), periodic mean removed, CG on
// -1/2 lap_nfh with the reference's avhs recurrence and stopping rule.
// Layout of one solve's buffers: b, r, p, ap, avhs (n each), then sc (16), the stencil /
// update partials (max(np, npu)), the update partials (npu), the dot partials
// (dot_nparts(n)), the boundary weights (3 x 9) and the state (2 ints), each 256-byte aligned.
namespace {
struct HartreeBufs {
  double *b, *r, *p, *ap, *avhs, *sc, *part, *upart, *dpart, *w;
  int *state;
  size_t bytes;
};
HartreeBufs hartree_layout(char *base, long n, int np, int npu) {
  HartreeBufs o{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *q = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return q;
  };
  o.b = (double *)take(sizeof(double) * n);
  o.r = (double *)take(sizeof(double) * n);
  o.p = (double *)take(sizeof(double) * n);
  o.ap = (double *)take(sizeof(double) * n);
  o.avhs = (double *)take(sizeof(double) * n);
  o.sc = (double *)take(sizeof(double) * 16);
  o.part = (double *)take(sizeof(double) * std::max(np, npu));
  o.upart = (double *)take(sizeof(double) * npu);
  o.dpart = (double *)take(sizeof(double) * dot_nparts(n));
  o.w = (double *)take(sizeof(double) * 27);
  o.state = (int *)take(sizeof(int) * 2);
  o.bytes = off;
  return o;
}
int hartree_update_parts(long n) {
  const long chunk = (long)param("hart_upd_chunk", 4096.0);
  return (int)std::min<long>(1024, std::max<long>(1, (n + chunk - 1) / chunk));
}
}  // namespace

size_t hartree_work_bytes(const Grid &g, int nfh) {
  const long n = g.n();
  return hartree_layout(nullptr, n, stencil_nparts(g, nfh, 1), hartree_update_parts(n)).bytes;
}

// The persistent CG (k_hcg_persistent) when the grid fits one cooperative launch of resident
// work-groups: returns the iterations, or 0 when it does not apply (the caller then runs the
// iteration blocks).  vh, b, p, avhs and sc[0] hold the warm start.
static int hartree_persistent(const HartreeArgs &a, const Grid &g, long n, const double *b, double *p,
                              const double *avhs, const double *sc, double dvo, int *state) {
  if (a.nfh < 4 || a.nfh > 6) return 0;
  if (a.nx % HCGP_T || a.ny % HCGP_T || a.nz % HCGP_P) return 0;
  const int bx = (a.nx + HCGP_T - 1) / HCGP_T, by = (a.ny + HCGP_T - 1) / HCGP_T,
            bz = (a.nz + HCGP_P - 1) / HCGP_P;
  const long G = (long)bx * by * bz;
  const void *fn = a.nfh == 4 ? (const void *)k_hcg_persistent<4>
                 : a.nfh == 5 ? (const void *)k_hcg_persistent<5> : (const void *)k_hcg_persistent<6>;
  static int coop = -1, ncu = 0;
  static int resident[3] = {-1, -1, -1};
  if (coop < 0) {
    HIPCK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, lib().device));
    HIPCK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, lib().device));
  }
  int &per_cu = resident[a.nfh - 4];
  if (per_cu < 0) HIPCK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0));
  if (!coop || G > (long)per_cu * ncu) return 0;
  hipStream_t s = lib().stream;
  HPArgs h{};
  h.nx = a.nx; h.ny = a.ny; h.nz = a.nz;
  h.px = g.nperi >= 1; h.py = g.nperi >= 2; h.pz = g.nperi >= 3;
  h.bx = bx; h.by = by;
  double w[9];
  fd_coefficients(a.nfh, w);
  for (int m = 0; m <= a.nfh; m++) {
    h.cx[m] = w[m] / g.dx / g.dx * (-0.5);
    h.cy[m] = w[m] / g.dy / g.dy * (-0.5);
    h.cz[m] = w[m] / g.dz / g.dz * (-0.5);
  }
  h.c0 = h.cx[0] + h.cy[0] + h.cz[0];
  h.vh = a.vh; h.p = p; h.avhs = avhs; h.b = b; h.sc = sc;
  h.slot1 = (HcgSlot *)scratch("hart.pslots", sizeof(HcgSlot) * 2 * G);
  h.slot2 = h.slot1 + G;
  h.dvo = dvo; h.eps = a.epsvh;
  h.ncgmin = a.ncgmin; h.ncgmax = a.ncgmax; h.ncgres = a.ncgres;
  h.state = state;
  h.bar = (HcgBar *)scratch("hart.pbar", sizeof(HcgBar) + sizeof(unsigned) * G);
  static_assert(sizeof(HcgSlot) == 16, "slot layout: value, generation, pad");
  hipLaunchKernelGGL(k_hcgp_init, dim3(1), dim3(256), 0, s, h.bar, (int)G, (unsigned *)h.slot1,
                     (int)(2 * G), state);
  HIPCK(hipGetLastError());
  void *args[] = {&h};
  HIPCK(hipLaunchCooperativeKernel(fn, dim3((unsigned)G), dim3(256), args, 0, s));
  HcgBar hb{};
  int hst[2] = {0, 0};
  HIPCK(hipMemcpyAsync(hst, state, sizeof(hst), hipMemcpyDeviceToHost, s));
  HIPCK(hipMemcpyAsync(&hb, h.bar, sizeof(hb), hipMemcpyDeviceToHost, s));
  sync();
  if (hst[0] != 1 || hb.err != 0)
    throw Error(1, "scf_hartree: a grid barrier of the persistent CG did not complete");
  return hst[1];
}

int hartree(const HartreeArgs &a, const HartreeWork *ws) {
  // with a caller workspace nothing is timed (events are library resources) and nothing is
  // allocated
  std::unique_ptr<Timed> ph;
  if (!ws) ph.reset(new Timed("phase_hartree", 0.0));
  hipStream_t s = lib().stream;
  Grid g{a.nx, a.ny, a.nz, a.nfh, a.nperi, 2.0 * a.xmax / a.nx, 2.0 * a.ymax / a.ny,
         2.0 * a.zmax / a.nz};
  const long n = g.n();
  const double omega = 8.0 * a.xmax * a.ymax * a.zmax;
  const int np = stencil_nparts(g, a.nfh, 1);
  // update work-groups: 4096 points each (512 at 128^3; 8192 points each, one work-group per
  // CU, left the update latency-bound: a 128^3 solve took 33.9 ms against 29.7)
  const int npu = hartree_update_parts(n);
  HartreeBufs B;
  if (ws) {
    B = hartree_layout(ws->base, n, np, npu);
    if (B.bytes > ws->bytes) throw Error(1, "scf_hartree: workspace too small for this grid");
  } else {
    B.b = (double *)scratch("hart.b", sizeof(double) * n);
    B.r = (double *)scratch("hart.r", sizeof(double) * n);
    B.p = (double *)scratch("hart.p", sizeof(double) * n);
    B.ap = (double *)scratch("hart.ap", sizeof(double) * n);
    B.avhs = (double *)scratch("hart.avhs", sizeof(double) * n);
    B.sc = (double *)scratch("hart.sc", sizeof(double) * 16);
    B.part = (double *)scratch("hart.part", sizeof(double) * std::max(np, npu));
    B.upart = (double *)scratch("hart.upart", sizeof(double) * npu);
    B.dpart = nullptr;  // launch_dots' own buffer
    B.w = (double *)scratch("hart.w", sizeof(double) * 27);
    B.state = (int *)scratch("hart.state", sizeof(int) * 2);
  }
  double *b = B.b, *r = B.r, *p = B.p, *ap = B.ap, *avhs = B.avhs, *sc = B.sc, *part = B.part;
  hipLaunchKernelGGL(k_hartree_b, dim3(nblk(n, 256)), dim3(256), 0, s, n, a.rho, b);
  double w[9];
  fd_coefficients(a.nfh, w);
  if (a.nd == 0) {
    const double hs[3] = {g.dx, g.dy, g.dz};
    const double *bd[3] = {a.boundx, a.boundy, a.boundz};
    const int need[3] = {a.nperi == 0, a.nperi <= 1, a.nperi <= 2};
    for (int axis = 2; axis >= 0; axis--) {
      if (!need[axis]) continue;
      // weights through a kernel argument copy: no host buffer has to outlive the launch
      double wc[9];
      for (int m = 0; m < 9; m++) wc[m] = w[m] / hs[axis] / hs[axis] * (-0.5);
      double *wd = B.w + 9 * axis;
      hipLaunchKernelGGL(k_set9, dim3(1), dim3(9), 0, s, Nine{{wc[0], wc[1], wc[2], wc[3], wc[4],
                         wc[5], wc[6], wc[7], wc[8]}}, wd);
      long faces = axis == 0 ? (long)a.ny * a.nz : axis == 1 ? (long)a.nx * a.nz : (long)a.nx * a.ny;
      hipLaunchKernelGGL(k_hartree_bound, dim3(nblk(faces, 256)), dim3(256), 0, s, a.nx, a.ny, a.nz,
                         a.nfh, axis, wd, bd[axis], b);
    }
  }
  if (a.nperi == 3) {
    double *ones = ap;  // ap is free until the warm-start product below
    launch_fill(n, 1.0, ones, s);
    launch_dots(n, b, n, ones, n, 1, g.dv() / omega, sc + 4, s, B.dpart);
    hipLaunchKernelGGL(k_sub_scalar, dim3(nblk(n, 256)), dim3(256), 0, s, n, sc + 4, 1.0, b);
  }
  // Av for the warm start
  StencilCall st;
  st.g = &g; st.in = a.vh; st.ld_in = n; st.out = ap; st.ld_out = n; st.veff = nullptr;
  st.nb = 1; st.nfh_coeffs = a.nfh;
  launch_stencil(st, s);
  HIPCK(hipMemcpyAsync(avhs, ap, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  launch_axpby(n, 1, nullptr, 1.0, b, n, nullptr, -1.0, ap, n, r, n, s);
  HIPCK(hipMemcpyAsync(p, r, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  // CG iterations run as blocks of `blk` on the device (scf_hartree8f.F90:611-1143): the
  // stopping rule is evaluated on the device after every iteration and later iterations of
  // the block do nothing, so the result is that of the host-driven loop; the host reads the
  // state once per block.  With param "hartree_graph" (default 1) a block is captured once
  // as a HIP graph and replayed.
  int *state = B.state;
  double *upart = B.upart;
  hipLaunchKernelGGL(k_hcg_reset, dim3(1), dim3(1), 0, s, state);
  launch_dots(n, r, n, r, n, 1, 1.0, sc + 0, s, B.dpart);  // r1r of iteration 1
  st.in = p; st.out = ap; st.dot_partial = part; st.skip = state;
  const double dvo = g.dv() / omega;
  lib().params["hartree_persistent_used"] = 0.0;
  // opt-in: at 128^3 it takes 47 us per iteration against 55 for the iteration blocks, at
  // 96^3 and 64^3 45 / 38 against 43 / 34 (its iterations are bound by memory round trips)
  if (!ws && param("hartree_persistent", 0.0) >= 1.0) {
    const int loops = hartree_persistent(a, g, n, b, p, avhs, sc, dvo, state);
    if (loops > 0) {
      lib().params["hartree_persistent_used"] = 1.0;
      return loops;
    }
  }
  const int blk = std::max(1, (int)param("hartree_block", 16.0));
  auto iteration = [&]() {
    launch_stencil(st, s);
    hipLaunchKernelGGL(k_hcg_update, dim3(npu), dim3(256), 0, s, n, sc, a.vh, p, ap, avhs, b, r,
                       part, np, upart, npu, state);
    hipLaunchKernelGGL(k_hcg_dir, dim3(std::min(nblk(n, 256), 2048u)), dim3(256), 0, s, n, sc, upart, npu,
                       a.ncgres, r, p, state);
    hipLaunchKernelGGL(k_hcg_ctl, dim3(1), dim3(1), 0, s, sc, dvo, a.epsvh, a.ncgmin, a.ncgmax,
                       state);
  };
  // a captured block is valid for the same buffers, sizes and stopping parameters
  struct Key {
    std::vector<const void *> ptrs;
    std::vector<long> ints;
    std::vector<double> reals;
    bool operator==(const Key &o) const {
      return ptrs == o.ptrs && ints == o.ints && reals == o.reals;
    }
  };
  struct GraphCache {
    hipGraphExec_t exec = nullptr;
    Key key;
  };
  static GraphCache gc;
  const bool use_graph = !ws && param("hartree_graph", 1.0) >= 1.0;
  if (use_graph) {
    const Key key{{a.vh, b, r, p, ap, avhs, sc, part, upart, state},
                  {n, a.nx, a.ny, a.nz, a.nperi, a.nfh, blk, a.ncgres, a.ncgmin, a.ncgmax, np, npu},
                  {a.epsvh, dvo, g.dx, g.dy, g.dz}};
    if (!(gc.exec && gc.key == key)) {
      if (gc.exec) HIPCK(hipGraphExecDestroy(gc.exec));
      gc.exec = nullptr;
      hipGraph_t graph;
      lib().capturing = true;
      try {
        HIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        for (int it = 0; it < blk; it++) iteration();
        HIPCK(hipStreamEndCapture(s, &graph));
      } catch (...) {
        lib().capturing = false;
        hipGraph_t g2;
        (void)hipStreamEndCapture(s, &g2);
        throw;
      }
      lib().capturing = false;
      HIPCK(hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0));
      HIPCK(hipGraphDestroy(graph));
      gc.key = key;
    }
  }
  int hst[2] = {0, 0};
  for (;;) {
    if (use_graph) HIPCK(hipGraphLaunch(gc.exec, s));
    else
      for (int it = 0; it < blk; it++) iteration();
    HIPCK(hipMemcpyAsync(hst, state, sizeof(hst), hipMemcpyDeviceToHost, s));
    sync();
    if (hst[0]) break;
    if (hst[1] > a.ncgmax) break;  // not reached: the stopping rule sets done at ncgmax
  }
  return hst[1];
}

void correct_potential(long n, double dvol_over_omega, double *vh) {
  hipStream_t s = lib().stream;
  double *sc = (double *)scratch("corr.sc", sizeof(double) * 2);
  double *ones = (double *)scratch("corr.ones", sizeof(double) * n);
  launch_fill(n, 1.0, ones, s);
  launch_dots(n, vh, n, ones, n, 1, dvol_over_omega, sc, s);
  hipLaunchKernelGGL(k_sub_scalar, dim3(nblk(n, 256)), dim3(256), 0, s, n, sc, 1.0, vh);
}

void charge(long n, int M, const double *psi, const double *f, double *rho) {
  hipStream_t s = lib().stream;
  const int chunk = 64;
  for (int l0 = 0; l0 < M; l0 += chunk) {
    const int nb = std::min(chunk, M - l0);
    hipLaunchKernelGGL(k_charge, dim3(nblk(n, 256)), dim3(256), 0, s, n, nb, psi + n * l0, n,
                       f + l0, rho);
  }
  HIPCK(hipGetLastError());
}

void force_eig(Ctx &c, int M, int ns, const double *psi, const double *sval, const double *fnele,
               const int *is_in_dev, double *fa, int nf) {
  const long n = c.g.n(), n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
  double w[9] = {0};
  switch (nf) {
    case 1: w[1] = 0.5; break;
    case 2: w[1] = 8.0 / 12.0; w[2] = -1.0 / 12.0; break;
    case 3: w[1] = 45.0 / 60.0; w[2] = -9.0 / 60.0; w[3] = 1.0 / 60.0; break;
    case 4: w[1] = 672.0 / 840.0; w[2] = -168.0 / 840.0; w[3] = 32.0 / 840.0; w[4] = -3.0 / 840.0; break;
    default: throw Error(7, "force_eig_r: nf outside 1-4 is not part of this build");
  }
  double *wd = c.buf("force.w", 9);
  HIPCK(hipMemcpyAsync(wd, w, sizeof(w), hipMemcpyHostToDevice, c.s));
  double *rp = c.buf("force.rp", (size_t)std::max<long>(1, n1) * M);
  c.proj(psi, n, M, rp, n1);
  const int chunk = std::max(1, std::min(M, 64));
  double *dpsi = c.buf("force.dpsi", (size_t)n * chunk);
  double *rd = c.buf("force.rd", (size_t)std::max<long>(1, n1) * chunk);
  const double hs[3] = {c.g.dx, c.g.dy, c.g.dz};
  for (int ax = 0; ax < 3; ax++) {
    for (int l0 = 0; l0 < M; l0 += chunk) {
      const int nb = std::min(chunk, M - l0);
      hipLaunchKernelGGL(k_fd1, dim3(nblk(n, 256), nb), dim3(256), 0, c.s, c.g.nx, c.g.ny, c.g.nz,
                         c.g.nperi >= 1, c.g.nperi >= 2, c.g.nperi >= 3, ax, nf, wd, 1.0 / hs[ax],
                         psi + n * l0, n, dpsi, n);
      c.proj(dpsi, n, nb, rd, n1);
      hipLaunchKernelGGL(k_force_acc, dim3(nblk(c.pl.nact, 64)), dim3(64), 0, c.s, c.pl.act,
                         c.pl.nact, is_in_dev, c.pl.nprjmx, c.pl.nums, ns, c.dij, c.sss,
                         rp + n1 * l0, rd, n1, nb, fnele + l0, sval + l0, fa + (long)ax * c.pl.natom);
    }
  }
  HIPCK(hipGetLastError());
}

}  // namespace kukan

namespace kukan {
// ============================================================================= differentiation
void band_energy(Ctx &c, int M, const double *veff, const double *psi, double *eband) {
  const long n = c.g.n(), n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
  double *rp = c.buf("adj.rp", (size_t)std::max<long>(1, n1) * M);
  double *hp = c.buf("tmp.nM", (size_t)n * M);
  c.proj(psi, n, M, rp, n1);
  c.hpsi(1, veff, psi, n, M, rp, n1, hp, n, false, eband, true);
}

void energy_adjoint(Ctx &c, int M, const double *psi, const double *f, const double *dvnl,
                    const int *is_in_dev, double *dedv, double *dedr) {
  const long n = c.g.n(), n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
  // dE/dveff(r) = dV sum_l f_l psi_l(r)^2
  launch_fill(n, 0.0, dedv, c.s);
  charge(n, M, psi, f, dedv);
  hipLaunchKernelGGL(k_scal, dim3(nblk(n, 256)), dim3(256), 0, c.s, n, c.g.dv(), dedv);
  launch_fill(3L * c.pl.natom, 0.0, dedr, c.s);
  if (c.pl.nact == 0 || n1 == 0) return;
  double *rp = c.buf("adj.rp", (size_t)n1 * M);
  double *drp = c.buf("adj.drp", (size_t)3 * n1 * M);
  c.proj(psi, n, M, rp, n1);
  const long tab = (long)c.pl.num_list * c.pl.nprjmx * c.pl.num_ppcell;
  for (int d = 0; d < 3; d++) {
    NonlocalPlan pd = c.pl;  // same lists, derivative table
    pd.vnl = dvnl + d * tab;
    launch_gather(pd, psi, n, M, drp + (long)d * n1 * M, n1, false, c.s);
  }
  hipLaunchKernelGGL(k_adj_force, dim3(c.pl.nact, 3), dim3(256), 0, c.s, c.pl.act, c.pl.nprjmx,
                     c.pl.nums, c.dij, rp, drp, n1, M, f, is_in_dev, c.pl.natom, dedr);
  HIPCK(hipGetLastError());
}
}  // namespace kukan
