// bloch.hpp — the complex (Bloch, k-point) path's context and solvers (bloch.hip): nrc = 1,
// collinear spin, no spin-orbit.  Complex arrays are double2 (complex*16: re, im).
#pragma once
#include <complex>

#include "kukan_internal.hpp"
#include "solvers.hpp"

namespace kukan {

// Per-call state of the complex path: the real path's context (grid, projector plan, dij/sss,
// stream) plus k and the projector phase table e^{i k.x} of every list entry.
struct ZCtx {
  Ctx &c;
  double kx, ky, kz;
  double2 *ph = nullptr;
  int ns_h = 1;  // spin index of dij in H
  ZCtx(Ctx &c_, double kx_, double ky_, double kz_) : c(c_), kx(kx_), ky(ky_), kz(kz_) {}
  long n() const { return c.g.n(); }
  long n1() const { return (long)c.pl.nprjmx * c.pl.num_ppcell; }
  double2 *zbuf(const char *tag, size_t count);
  // the phase table from the list offsets and atom indices (device int arrays)
  void phases(const int *lstx, const int *lsty, const int *lstz, const int *natx, const int *naty,
              const int *natz);
  // cp(:, b) = <p|u_b> (+= with accumulate)
  void proj(const double2 *u, long ldu, int nb, double2 *cp, long ldcp, bool accumulate = false);
  // out_b += conj(P) X cp_b / dV; X = dij (nhs = 1, spin ns) or sss (nhs = 0)
  void expand(int nhs, int ns, const double2 *cp, long ldcp, int nb, double2 *out, long ldo);
  // out_b = H_k u_b with the given <p|u_b>
  void hpsi(const double *veff, const double2 *u, long ldu, int nb, const double2 *cp, long ldcp,
            double2 *out, long ldo);
  // out_b = S u_b
  void spsi(const double2 *u, long ldu, int nb, const double2 *cp, long ldcp, double2 *out,
            long ldo);
  // out_b = K u_b for nb bands (band strides default to n)
  void precond(const double2 *u, double2 *out, int nb = 1, long ldu = 0, long ldo = 0);
  // out_b = ca[b] x_b + cb[b] y_b with host coefficient arrays (y may be null)
  void comb(long len, int nb, const std::complex<double> *ca, const double2 *x, long ldx,
            const std::complex<double> *cb, const double2 *y, long ldy, double2 *out, long ldo);
  // res_b = scale sum conj(x_b) y_b (device)
  void dots(const double2 *x, long ldx, const double2 *y, long ldy, int nb, double2 *res,
            double scale);
  std::complex<double> dot1(const double2 *x, const double2 *y, double scale);
  void axpby(long len, std::complex<double> a, const double2 *x, std::complex<double> b,
             const double2 *y, double2 *out);
  void scale(long len, double s, double2 *x);
};

// orthogonalization_c_01 (bands [l_inp - 1] or all when l_inp = 0)
void zortho_c_01(ZCtx &z, int M, int l_inp, int nflag, int key_cmpt, double2 *psi, double2 *spsi,
                 double2 *cps);
void zrayleigh_ritz(ZCtx &z, int M, const double *veff, bool l_rrb, double2 *psi, double2 *spsi,
                    double2 *cps, double2 *hpsi, double *sval);
void zkscg_c(ZCtx &z, int M, int northo, int nkscg, int nsdmax, int nprecon_cg, int key_cmpt,
             int key_nocmpt, double epssd, const double *veff, double2 *Psi, double2 *SPsi,
             double2 *Cps, double2 *H, double *sval, double *resid, int *ksconv, int *ksitmax);
// ksdiis_c for one spin / k-point (ksdiis8f.F90:715-1509)
struct ZDiisArgs {
  int M, northo, ndiismax, nprecon_diis, key_cmpt;
  bool l_hsv;
  double eps_eig_diis, alambda_diis, alambda_min, alambda_max, epssd;
};
void zksdiis_c(ZCtx &z, const ZDiisArgs &a, const double *veff, double2 *Psi, double2 *SPsi,
               double2 *Cps, double2 *H, double *sval, double *resid, int *ksconv, int *ksitmax);
void zreorder(ZCtx &z, int M, double *sval, double *resid, double2 *psi, double2 *spsi,
              double2 *cps, double2 *hpsi);
void zcharge(long n, int nb, const void *u, long ldu, const double *f, double *rho, hipStream_t s);
void zhalo_fill(int nperi, int na, int nb, int nc, int nf, int nf1, int nsnd, int ncol, void *v,
                hipStream_t s);
void zlap_leaf(const Grid &g, int nf, double xmax, double ymax, double zmax, double kx, double ky,
               double kz, void *vcm, const double *veff, void *avcm_l, void *workc, hipStream_t s);
void zprecond_leaf(const Grid &g, int nf, const void *vcm, void *avcm, int ncol, hipStream_t s);

// scf_diag's arguments for its complex branch, with the step switches already decided
struct BlochDiagArgs {
  int nso, ncol, iscf, ncgscf, nretcg, natom, num_spe, neigmx, nums, nspv, numk, nperi, northo, nf;
  int num_list, num_ppcell, nprjmx, ncpx, ncpy, ncpz, key_cmpt, key_nocmpt, key_in, key_inps;
  int nkscg, nsdmax, nprecon_cg, ndiismax, nprecon_diis;
  bool l_rrb, l_rra, l_reo, l_hsv;
  double dx, dy, dz, epssd, eps_eig_diis, alambda_diis, alambda_min, alambda_max;
  const double *skpxx, *skpyy, *skpzz, *dij, *veff, *vnlocp, *sss;
  const int *indspe, *natpri, *naps, *nprj, *natinf, *lstvec2, *lstx, *lsty, *lstz, *natx, *naty,
      *natz;
  void *cspsep, *sveccm, *hsvcm, *ssvcm;
  double *sval, *residual_states;
  int *ksconv, *ksitmax;
};
void scf_diag_bloch(const BlochDiagArgs &a);
void reorder_bloch(int num_ppcell, int nprjmx, int neigmx, int nums, int ncol, int numk, int ncpx,
                   int ncpy, int ncpz, int ns1, int nk, double *sval, double *residual_states,
                   void *cspsep, void *sveccm, void *ssvcm, void *hsvcm);

}  // namespace kukan
