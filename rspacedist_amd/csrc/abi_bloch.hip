// abi_bloch.hip — extern "C" entry points of the complex (Bloch, k-point) path
// (declared in include/kukan_hip.h): nrc = 1, collinear spin (ncol = 1), no spin-orbit (nso = 0).
// Complex arrays are complex*16 (re, im pairs of doubles) in the reference's column-major layouts;
// host or device pointers as everywhere else in the library.
#include <algorithm>

#include "abi_common.hpp"
#include "bloch.hpp"
#include "kukan_hip.h"

using namespace kukan;
using namespace kukan::abi;

namespace {

void need_collinear(int ncol, int nso, const char *who) {
  if (ncol != 1 || nso != 0)
    throw Error(5, std::string(who) + ": noncollinear spin and spin-orbit (ncol = 2, nso = 1) are not part "
                                      "of this library");
}

// The device copies of the projector tables shared by the nonlocal, orthogonalisation and
// solver entry points, the plan built from them and the phase table at one k-point.
struct ZCall {
  Ctx c;
  Arr vnl, lst, lx, ly, lz, ax, ay, az;
  ZCtx z;
  ZCall(const Grid &g, double kx, double ky, double kz) : c(g), z(c, kx, ky, kz) {}
  void plan(int natom, int nprjmx, int num_spe, int num_ppcell, int num_list, int nums, int key_in,
            int key_inps, const int *nprj, const int *indspe, const int *natinf, const int *natpri,
            const int *naps, const double *vnlocp, const int *lstvec2, const int *lstx,
            const int *lsty, const int *lstz, const int *natx, const int *naty, const int *natz,
            const double *dij, const double *sss_dev, const double *sss_host) {
    const size_t nl = (size_t)num_list * num_ppcell;
    vnl = Arr(vnlocp, nl * nprjmx * D, Arr::In);
    lst = Arr(lstvec2, nl * I, Arr::In);
    lx = Arr(lstx, nl * I, Arr::In);
    ly = Arr(lsty, nl * I, Arr::In);
    lz = Arr(lstz, nl * I, Arr::In);
    ax = Arr(natx, (size_t)natom * I, Arr::In);
    ay = Arr(naty, (size_t)natom * I, Arr::In);
    az = Arr(natz, (size_t)natom * I, Arr::In);
    setup(c, natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_in, key_inps, nprj, indspe,
          natinf, natpri, naps, vnl, lst, dij, sss_dev, sss_host);
    z.phases(lx.i(), ly.i(), lz.i(), ax.i(), ay.i(), az.i());
  }
};

inline size_t kofs(long len, int neigmx, int nums, int ns, int nk) {
  return (size_t)len * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
}

}  // namespace

extern "C" {

int kukan_overlap_finitedifference_c(int nperi, int ncpa, int ncpb, int ncpc, int nf, int nf1,
                                     int nsnd, int ncol, void *v_cc) {
  API_BEGIN
  const size_t cnt = (size_t)(ncpa + nf + nf1 + 1) * (ncpb + nf + nf1 + 1) * (ncpc + nf + nf1 + 1) * ncol;
  Arr v(v_cc, cnt * Z, Arr::InOut);
  zhalo_fill(nperi, ncpa, ncpb, ncpc, nf, nf1, nsnd, ncol, v.v(), lib().stream);
  v.finish();
  API_END
}

int kukan_kslaplacian_c(int ncpx, int ncpy, int ncpz, int neigmx, int ncol, int nvef, int nf,
                        double dx, double dy, double dz, double xmax, double ymax, double zmax,
                        double skpx, double skpy, double skpz, void *vcm, const double *veff,
                        int l, void *avcm, void *workc) {
  API_BEGIN
  need_collinear(ncol, 0, "kslaplacian_c");
  if (nf < 1 || nf > 8) throw Error(1, "error in kslaplacian_c! This subroutine is for nf in 1-8.");
  Grid g{ncpx, ncpy, ncpz, nf, 3, dx, dy, dz};
  const long n = g.n();
  Arr in(vcm, (size_t)(ncpx + 2 * nf) * (ncpy + 2 * nf) * (ncpz + 2 * nf) * Z, Arr::InOut);
  Arr ve(veff, (size_t)n * std::max(1, nvef) * D, Arr::In);
  Arr out(avcm, (size_t)n * neigmx * Z, Arr::InOut);
  Arr w(workc, (size_t)n * Z, Arr::InOut);
  zlap_leaf(g, nf, xmax, ymax, zmax, skpx, skpy, skpz, in.v(), ve.d(),
            (double2 *)out.v() + (size_t)(l - 1) * n, w.v(), lib().stream);
  in.finish(); out.finish(); w.finish();
  API_END
}

int kukan_ksprecondition_c(int ncpx, int ncpy, int ncpz, int nf, int ncol, const void *vcm,
                           void *avcm) {
  API_BEGIN
  if (nf < 1) throw Error(1, "ksprecondition_c: the halo width nf must be at least 1");
  Grid g{ncpx, ncpy, ncpz, nf, 3, 1.0, 1.0, 1.0};
  const long nh = (long)(ncpx + 2 * nf) * (ncpy + 2 * nf) * (ncpz + 2 * nf);
  Arr in(vcm, (size_t)nh * ncol * Z, Arr::In);
  Arr out(avcm, (size_t)g.n() * ncol * Z, Arr::Out);
  zprecond_leaf(g, nf, in.v(), out.v(), ncol, lib().stream);
  out.finish();
  API_END
}

int kukan_nonlocaloperation_c_01(int natom, int nprjmx, int num_spe, int num_ppcell, int num_list,
                                 int neigmx, int ncol, int l, int ncpx, int ncpy, int ncpz,
                                 int key_natpri_in, int key_natpri_inps, double dx, double dy,
                                 double dz, double skpx, double skpy, double skpz, const int *nprj,
                                 const int *indspe, const int *natinf, const int *natpri,
                                 const int *naps, const int *lstvec2, const int *lstx,
                                 const int *lsty, const int *lstz, const int *natx,
                                 const int *naty, const int *natz, const double *vnlocp,
                                 const void *vcm2, void *cpsep, void *vcccm) {
  (void)vcccm;  // the reference's work array for the phased gathered values
  API_BEGIN
  need_collinear(ncol, 0, "nonlocaloperation_c_01");
  Grid g{ncpx, ncpy, ncpz, 1, 3, dx, dy, dz};
  ZCall zc(g, skpx, skpy, skpz);
  const long n = g.n();
  Arr v(vcm2, (size_t)n * neigmx * Z, Arr::In);
  Arr cp(cpsep, (size_t)nprjmx * num_ppcell * Z, Arr::InOut);
  zc.plan(natom, nprjmx, num_spe, num_ppcell, num_list, 1, key_natpri_in, key_natpri_inps, nprj,
          indspe, natinf, natpri, naps, vnlocp, lstvec2, lstx, lsty, lstz, natx, naty, natz, nullptr,
          nullptr, nullptr);
  zc.z.proj((const double2 *)v.v() + (size_t)(l - 1) * n, n, 1, (double2 *)cp.v(),
            (long)nprjmx * num_ppcell, true);
  cp.finish();
  API_END
}

int kukan_nonlocaloperation_c_02(int nhs, int nso, int natom, int num_spe, int neigmx, int nums,
                                 int nprjmx, int ncol, int ns1, int l, int num_list,
                                 int num_ppcell, int ncpx, int ncpy, int ncpz, int npxmax,
                                 int npymax, int npzmax, int key_natpri_in, int key_natpri_inps,
                                 int key_soc_calc, double dx, double dy, double dz, double skpx,
                                 double skpy, double skpz, const int *nprj, const int *indspe,
                                 const int *natinf, const int *lstvec2, const int *natpri,
                                 const int *naps, const int *natsoc, const int *lstx,
                                 const int *lsty, const int *lstz, const int *natx,
                                 const int *naty, const int *natz, const double *vnlocp,
                                 const void *cpsep, const double *dijsss, const double *dijsoc,
                                 void *avcm, void *vcccm, void *avc) {
  (void)npxmax; (void)npymax; (void)npzmax; (void)key_soc_calc; (void)natsoc; (void)dijsoc;
  (void)vcccm; (void)avc;  // the reference's work arrays
  API_BEGIN
  if (nhs > 1 || nso > 1 || nhs < 0 || nso < 0) throw Error(1, "error occurs in nonlocaloperation_c_02");
  if (nhs == 0 && ns1 != 1) throw Error(1, "nonlocaloperation_c_02: nhs=0 and ns1/=1");
  need_collinear(ncol, nso, "nonlocaloperation_c_02");
  Grid g{ncpx, ncpy, ncpz, 1, 3, dx, dy, dz};
  ZCall zc(g, skpx, skpy, skpz);
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell;
  const size_t xcnt = (size_t)nprjmx * nprjmx * (nums * nhs - nhs + 1) * (natom * nhs + num_spe * (1 - nhs));
  Arr cp(cpsep, (size_t)n1 * Z, Arr::In);
  Arr X(dijsss, xcnt * D, Arr::In);
  Arr out(avcm, (size_t)n * neigmx * Z, Arr::InOut);
  zc.plan(natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps, nprj,
          indspe, natinf, natpri, naps, vnlocp, lstvec2, lstx, lsty, lstz, natx, naty, natz,
          nhs == 1 ? X.d() : nullptr, nhs == 0 ? X.d() : nullptr, nullptr);
  zc.z.expand(nhs, ns1, (const double2 *)cp.v(), n1, 1, (double2 *)out.v() + (size_t)(l - 1) * n, n);
  out.finish();
  API_END
}

int kukan_orthogonalization_c_01(int natom, int neigmx, int nprjmx, int nums, int ncol, int numk,
                                 int nk, int ns1, int l_inp, int num_list, int num_ppcell,
                                 int num_spe, int nflag, int ndisp, int ncpx, int ncpy, int ncpz,
                                 int npxmax, int npymax, int npzmax,
                                 int key_ortho_cmpt_innerproduct, int key_natpri_in,
                                 int key_natpri_inps, double dx, double dy, double dz,
                                 double skpx, double skpy, double skpz, const int *nprj,
                                 const int *indspe, const int *natpri, const int *naps,
                                 const int *natinf, const int *lstvec2, const int *latom,
                                 const int *lstx, const int *lsty, const int *lstz,
                                 const int *natx, const int *naty, const int *natz,
                                 const double *vnlocp, const double *sss, void *sveccm,
                                 void *ssvcm, void *cspsep) {
  (void)ndisp; (void)npxmax; (void)npymax; (void)npzmax; (void)latom;
  API_BEGIN
  need_collinear(ncol, 0, "orthogonalization_c_01");
  Grid g{ncpx, ncpy, ncpz, 1, 3, dx, dy, dz};
  ZCall zc(g, skpx, skpy, skpz);
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  Arr Sss(sss, (size_t)nprjmx * nprjmx * num_spe * D, Arr::In);
  Arr Ps(sveccm, (size_t)n * nall * Z, Arr::InOut);
  Arr Sp(ssvcm, (size_t)n * nall * Z, Arr::InOut);
  Arr Cs(cspsep, (size_t)n1 * nall * Z, Arr::InOut);
  zc.plan(natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps, nprj,
          indspe, natinf, natpri, naps, vnlocp, lstvec2, lstx, lsty, lstz, natx, naty, natz, nullptr,
          Sss.d(), sss);
  zortho_c_01(zc.z, neigmx, l_inp, nflag, key_ortho_cmpt_innerproduct,
              (double2 *)Ps.v() + kofs(n, neigmx, nums, ns1, nk),
              (double2 *)Sp.v() + kofs(n, neigmx, nums, ns1, nk),
              (double2 *)Cs.v() + kofs(n1, neigmx, nums, ns1, nk));
  Ps.finish(); Sp.finish(); Cs.finish();
  API_END
}

int kukan_orthogonalization_c(int natom, int neigmx, int nprjmx, int nums, int ncol, int numk,
                              int num_ppcell, int num_spe, int northo, int num_list, int nflag,
                              int ndisp, int ncpx, int ncpy, int ncpz, int npxmax, int npymax,
                              int npzmax, int key_ortho_cmpt_innerproduct, int key_natpri_in,
                              int key_natpri_inps, double dx, double dy, double dz,
                              const int *nprj, const int *indspe, const int *natpri,
                              const int *naps, const int *natinf, const int *lstvec2,
                              const int *latom, const int *ntyppp, const int *lstx,
                              const int *lsty, const int *lstz, const int *natx, const int *naty,
                              const int *natz, const double *skpxx, const double *skpyy,
                              const double *skpzz, const double *vnlocp, const double *sss,
                              void *sveccm, void *ssvcm, void *cspsep) {
  (void)ndisp; (void)npxmax; (void)npymax; (void)npzmax; (void)latom; (void)ntyppp;
  API_BEGIN
  need_collinear(ncol, 0, "orthogonalization_c");
  Grid g{ncpx, ncpy, ncpz, 1, 3, dx, dy, dz};
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  const std::vector<double> kx = host_doubles(skpxx, numk), ky = host_doubles(skpyy, numk),
                            kz = host_doubles(skpzz, numk);
  Arr Sss(sss, (size_t)nprjmx * nprjmx * num_spe * D, Arr::In);
  Arr Ps(sveccm, (size_t)n * nall * Z, Arr::InOut);
  Arr Sp(ssvcm, (size_t)n * nall * Z, Arr::InOut);
  Arr Cs(cspsep, (size_t)n1 * nall * Z, Arr::InOut);
  for (int nk = 1; nk <= numk; nk++) {
    ZCall zc(g, kx[nk - 1], ky[nk - 1], kz[nk - 1]);
    zc.plan(natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps, nprj,
            indspe, natinf, natpri, naps, vnlocp, lstvec2, lstx, lsty, lstz, natx, naty, natz,
            nullptr, Sss.d(), sss);
    for (int ns = 1; ns <= nums; ns++)
      for (int loop = 0; loop < northo; loop++)
        zortho_c_01(zc.z, neigmx, 0, nflag, key_ortho_cmpt_innerproduct,
                    (double2 *)Ps.v() + kofs(n, neigmx, nums, ns, nk),
                    (double2 *)Sp.v() + kofs(n, neigmx, nums, ns, nk),
                    (double2 *)Cs.v() + kofs(n1, neigmx, nums, ns, nk));
  }
  Ps.finish(); Sp.finish(); Cs.finish();
  API_END
}

int kukan_kscg_c(int nso, int natom, int num_spe, int neigmx, int nums, int ncol, int nspv,
                 int numk, int nperi, int northo, int ns1, int nk, int nf, int nkscg, int nsdmax,
                 int nprecon_cg, int nprjmx, int num_list, int num_ppcell, int ndisp, int ncpx,
                 int ncpy, int ncpz, int npxmax, int npymax, int npzmax,
                 int key_ortho_nocmpt_innerproduct, int key_ortho_cmpt_innerproduct,
                 int key_natpri_in, int key_natpri_inps, int key_natpri_out, int key_soc_calc,
                 double dx, double dy, double dz, double skpx, double skpy, double skpz,
                 double epssd, const int *indspe, const int *natpri, const int *naps,
                 const int *nprj, const int *natinf, const int *lstvec2, const int *latom,
                 const int *natsoc, const int *ntyppp, const int *lstx, const int *lsty,
                 const int *lstz, const int *natx, const int *naty, const int *natz,
                 const double *dij, const double *dijsoc, const double *veff,
                 const double *vnlocp, const double *sss, int *ksconv, int *ksitmax,
                 double *sval, void *cspsep, void *sveccm, void *ssvcm,
                 double *residual_states, void *hsvcm) {
  (void)ndisp; (void)npxmax; (void)npymax; (void)npzmax; (void)key_natpri_out;
  (void)key_soc_calc; (void)latom; (void)natsoc; (void)ntyppp; (void)dijsoc;
  API_BEGIN
  need_collinear(ncol, nso, "kscg_c");
  Grid g{ncpx, ncpy, ncpz, nf, nperi, dx, dy, dz};
  ZCall zc(g, skpx, skpy, skpz);
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  Arr Dij(dij, (size_t)nprjmx * nprjmx * nums * natom * D, Arr::In);
  Arr Sss(sss, (size_t)nprjmx * nprjmx * num_spe * D, Arr::In);
  Arr Ve(veff, (size_t)n * nspv * D, Arr::In);
  Arr Sv(sval, (size_t)nall * D, Arr::InOut);
  Arr Cs(cspsep, (size_t)n1 * nall * Z, Arr::InOut);
  Arr Ps(sveccm, (size_t)n * nall * Z, Arr::InOut);
  Arr Sp(ssvcm, (size_t)n * nall * Z, Arr::InOut);
  Arr Re(residual_states, (size_t)neigmx * D, Arr::InOut);
  Arr Hp(hsvcm, (size_t)n * neigmx * Z, Arr::InOut);
  zc.plan(natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps, nprj,
          indspe, natinf, natpri, naps, vnlocp, lstvec2, lstx, lsty, lstz, natx, naty, natz,
          Dij.d(), Sss.d(), sss);
  zc.z.ns_h = ns1;
  const int nsv = std::min(ns1, nspv);
  zkscg_c(zc.z, neigmx, northo, nkscg, nsdmax, nprecon_cg, key_ortho_cmpt_innerproduct,
          key_ortho_nocmpt_innerproduct, epssd, Ve.d() + (size_t)n * (nsv - 1),
          (double2 *)Ps.v() + kofs(n, neigmx, nums, ns1, nk),
          (double2 *)Sp.v() + kofs(n, neigmx, nums, ns1, nk),
          (double2 *)Cs.v() + kofs(n1, neigmx, nums, ns1, nk), (double2 *)Hp.v(),
          Sv.d() + kofs(1, neigmx, nums, ns1, nk), Re.d(), ksconv, ksitmax);
  Sv.finish(); Cs.finish(); Ps.finish(); Sp.finish(); Re.finish(); Hp.finish();
  API_END
}

int kukan_ksdiis_c(int l_hsv, int nso, int natom, int num_spe, int neigmx, int nums, int ncol,
                   int nspv, int numk, int nperi, int ndiismax, int nprecon_diis, int nk, int ns1,
                   int nf, int nprjmx, int num_list, int num_ppcell, int ndisp, int northo, int ncpx,
                   int ncpy, int ncpz, int npxmax, int npymax, int npzmax,
                   int key_ortho_cmpt_innerproduct, int key_natpri_in, int key_natpri_inps,
                   int key_soc_calc, double eps_eig_diis, double alambda_diis, double ratio_diis,
                   double alambda_min, double alambda_max, double epssd, double xmax, double ymax,
                   double zmax, double skpx, double skpy, double skpz, const int *indspe,
                   const int *natpri, const int *naps, const int *nprj, const int *natinf,
                   const int *lstvec2, const int *latom, const int *natsoc, const int *ntyppp,
                   const int *lstx, const int *lsty, const int *lstz, const int *natx,
                   const int *naty, const int *natz, const double *dij, const double *dijsoc,
                   const double *veff, const double *vnlocp, const double *sss, int *ksconv,
                   int *ksitmax, void *cspsep, double *sval, void *sveccm, void *ssvcm,
                   double *residual_states, void *hsvcm) {
  (void)ndisp; (void)npxmax; (void)npymax; (void)npzmax; (void)key_soc_calc; (void)ratio_diis;
  (void)latom; (void)natsoc; (void)ntyppp; (void)dijsoc;
  API_BEGIN
  need_collinear(ncol, nso, "ksdiis_c");
  // the grid spacing follows from the cell, as in the reference (ksdiis8f.F90:812-814)
  Grid g{ncpx, ncpy, ncpz, nf, nperi, 2.0 * xmax / ncpx, 2.0 * ymax / ncpy, 2.0 * zmax / ncpz};
  ZCall zc(g, skpx, skpy, skpz);
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  Arr Dij(dij, (size_t)nprjmx * nprjmx * nums * natom * D, Arr::In);
  Arr Sss(sss, (size_t)nprjmx * nprjmx * num_spe * D, Arr::In);
  Arr Ve(veff, (size_t)n * nspv * D, Arr::In);
  Arr Cs(cspsep, (size_t)n1 * nall * Z, Arr::InOut);
  Arr Sv(sval, (size_t)nall * D, Arr::InOut);
  Arr Ps(sveccm, (size_t)n * nall * Z, Arr::InOut);
  Arr Sp(ssvcm, (size_t)n * nall * Z, Arr::InOut);
  Arr Re(residual_states, (size_t)neigmx * D, Arr::InOut);
  Arr Hp(hsvcm, (size_t)n * neigmx * Z, Arr::InOut);
  zc.plan(natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps, nprj,
          indspe, natinf, natpri, naps, vnlocp, lstvec2, lstx, lsty, lstz, natx, naty, natz,
          Dij.d(), Sss.d(), sss);
  zc.z.ns_h = ns1;
  const ZDiisArgs da{neigmx, northo, ndiismax, nprecon_diis, key_ortho_cmpt_innerproduct, l_hsv != 0,
                     eps_eig_diis, alambda_diis, alambda_min, alambda_max, epssd};
  zksdiis_c(zc.z, da, Ve.d() + (size_t)n * (std::min(ns1, nspv) - 1),
            (double2 *)Ps.v() + kofs(n, neigmx, nums, ns1, nk),
            (double2 *)Sp.v() + kofs(n, neigmx, nums, ns1, nk),
            (double2 *)Cs.v() + kofs(n1, neigmx, nums, ns1, nk), (double2 *)Hp.v(),
            Sv.d() + kofs(1, neigmx, nums, ns1, nk), Re.d(), ksconv, ksitmax);
  Cs.finish(); Sv.finish(); Ps.finish(); Sp.finish(); Re.finish(); Hp.finish();
  API_END
}

int kukan_rayleighritz_c(int nso, int natom, int num_ppcell, int num_spe, int nprjmx, int neigmx,
                         int numk, int nums, int ncol, int nspv, int num_list, int ncpx, int ncpy,
                         int ncpz, int npxmax, int npymax, int npzmax, int nperi, int nk, int ns1,
                         int nf, int ndisp, int key_natpri_in, int key_natpri_inps,
                         int key_soc_calc, double dx, double dy, double dz, double xmax,
                         double ymax, double zmax, const double *skpxx, const double *skpyy,
                         const double *skpzz, const int *nprj, const int *natinf,
                         const int *lstvec2, const int *indspe, const int *natpri,
                         const int *naps, const int *natsoc, const int *ntyppp, const int *lstx,
                         const int *lsty, const int *lstz, const int *natx, const int *naty,
                         const int *natz, const double *dij, const double *dijsoc,
                         const double *veff, const double *vnlocp, int l_rrb, void *cspsep,
                         void *sveccm, void *ssvcm, void *hsvcm, double *sval) {
  (void)npxmax; (void)npymax; (void)npzmax; (void)ndisp; (void)key_soc_calc; (void)xmax;
  (void)ymax; (void)zmax; (void)natsoc; (void)ntyppp; (void)dijsoc;
  API_BEGIN
  need_collinear(ncol, nso, "rayleighritz_c");
  Grid g{ncpx, ncpy, ncpz, nf, nperi, dx, dy, dz};
  const std::vector<double> kx = host_doubles(skpxx, numk), ky = host_doubles(skpyy, numk),
                            kz = host_doubles(skpzz, numk);
  ZCall zc(g, kx[nk - 1], ky[nk - 1], kz[nk - 1]);
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  Arr Dij(dij, (size_t)nprjmx * nprjmx * nums * natom * D, Arr::In);
  Arr Ve(veff, (size_t)n * nspv * D, Arr::In);
  Arr Cs(cspsep, (size_t)n1 * nall * Z, Arr::InOut);
  Arr Ps(sveccm, (size_t)n * nall * Z, Arr::InOut);
  Arr Sp(ssvcm, (size_t)n * nall * Z, Arr::InOut);
  Arr Hp(hsvcm, (size_t)n * neigmx * Z, Arr::InOut);
  Arr Sv(sval, (size_t)nall * D, Arr::InOut);
  zc.plan(natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps, nprj,
          indspe, natinf, natpri, naps, vnlocp, lstvec2, lstx, lsty, lstz, natx, naty, natz,
          Dij.d(), nullptr, nullptr);
  zc.c.q_zero = false;  // the reference rotates S Psi explicitly here (no sss argument)
  zc.z.ns_h = ns1;
  const int nsv = std::min(ns1, nspv);
  zrayleigh_ritz(zc.z, neigmx, Ve.d() + (size_t)n * (nsv - 1), l_rrb != 0,
                 (double2 *)Ps.v() + kofs(n, neigmx, nums, ns1, nk),
                 (double2 *)Sp.v() + kofs(n, neigmx, nums, ns1, nk),
                 (double2 *)Cs.v() + kofs(n1, neigmx, nums, ns1, nk), (double2 *)Hp.v(),
                 Sv.d() + kofs(1, neigmx, nums, ns1, nk));
  Cs.finish(); Ps.finish(); Sp.finish(); Hp.finish(); Sv.finish();
  API_END
}

int kukan_scf_charge_smt_c(int neigmx, int nums, int ncol, int nspv, int numk, int ncpx, int ncpy,
                           int ncpz, const double *fnele, const void *sveccm, double *rhosmt) {
  API_BEGIN
  need_collinear(ncol, 0, "scf_charge_smt_c");
  const long n = (long)ncpx * ncpy * ncpz, nall = (long)neigmx * nums * numk;
  Arr F(fnele, (size_t)nall * D, Arr::In);
  Arr Ps(sveccm, (size_t)n * nall * Z, Arr::In);
  if (nspv > 1 && nspv < nums) throw Error(1, "scf_charge_smt_c: nspv < nums");
  Arr Rho(rhosmt, (size_t)n * nspv * D, Arr::Out);
  launch_fill(n * nspv, 0.0, Rho.d(), lib().stream);
  for (int nk = 1; nk <= numk; nk++)
    for (int ns = 1; ns <= nums; ns++) {  // spins add into one density when nspv = 1
      const size_t b = kofs(1, neigmx, nums, ns, nk);
      zcharge(n, neigmx, (const double2 *)Ps.v() + (size_t)n * b, n, F.d() + b,
              Rho.d() + (size_t)n * (nspv > 1 ? ns - 1 : 0), lib().stream);
    }
  Rho.finish();
  API_END
}

}  // extern "C"

namespace kukan {

// scf_diag's complex branch (scf_diag8f.F90:213-312): per k-point and spin, Rayleigh-Ritz
// (l_rrb), band-sequential kscg_c or RMM-DIIS ksdiis_c, Rayleigh-Ritz (l_rra), then the
// reorder (l_reo).
void scf_diag_bloch(const BlochDiagArgs &a) {
  need_collinear(a.ncol, a.nso, "scf_diag");
  const bool use_cg = a.iscf <= a.ncgscf || a.iscf % a.nretcg == 0;
  const Grid g{a.ncpx, a.ncpy, a.ncpz, a.nf, a.nperi, a.dx, a.dy, a.dz};
  const long n = g.n(), n1 = (long)a.nprjmx * a.num_ppcell, nall = (long)a.neigmx * a.nums * a.numk;
  const std::vector<double> kx = host_doubles(a.skpxx, a.numk), ky = host_doubles(a.skpyy, a.numk),
                            kz = host_doubles(a.skpzz, a.numk);
  Arr Dij(a.dij, (size_t)a.nprjmx * a.nprjmx * a.nums * a.natom * D, Arr::In);
  Arr Sss(a.sss, (size_t)a.nprjmx * a.nprjmx * a.num_spe * D, Arr::In);
  Arr Ve(a.veff, (size_t)n * a.nspv * D, Arr::In);
  Arr Cs(a.cspsep, (size_t)n1 * nall * Z, Arr::InOut);
  Arr Ps(a.sveccm, (size_t)n * nall * Z, Arr::InOut);
  Arr Sp(a.ssvcm, (size_t)n * nall * Z, Arr::InOut);
  Arr Hp(a.hsvcm, (size_t)n * a.neigmx * Z, Arr::InOut);
  Arr Sv(a.sval, (size_t)nall * D, Arr::InOut);
  Arr Re(a.residual_states, (size_t)nall * D, Arr::InOut);
  *a.ksconv = 1;
  *a.ksitmax = 0;
  for (int nk = 1; nk <= a.numk; nk++) {
    ZCall zc(g, kx[nk - 1], ky[nk - 1], kz[nk - 1]);
    zc.plan(a.natom, a.nprjmx, a.num_spe, a.num_ppcell, a.num_list, a.nums, a.key_in, a.key_inps,
            a.nprj, a.indspe, a.natinf, a.natpri, a.naps, a.vnlocp, a.lstvec2, a.lstx, a.lsty,
            a.lstz, a.natx, a.naty, a.natz, Dij.d(), Sss.d(), a.sss);
    const bool qz = zc.c.q_zero;
    for (int ns = 1; ns <= a.nums; ns++) {
      double2 *psi = (double2 *)Ps.v() + kofs(n, a.neigmx, a.nums, ns, nk);
      double2 *spsi = (double2 *)Sp.v() + kofs(n, a.neigmx, a.nums, ns, nk);
      double2 *cps = (double2 *)Cs.v() + kofs(n1, a.neigmx, a.nums, ns, nk);
      double *sv = Sv.d() + kofs(1, a.neigmx, a.nums, ns, nk);
      double *re = Re.d() + kofs(1, a.neigmx, a.nums, ns, nk);
      const double *vs = Ve.d() + (size_t)n * (std::min(ns, a.nspv) - 1);
      zc.z.ns_h = ns;
      auto rr = [&](bool l_rrb) {
        zc.c.q_zero = false;  // rayleighritz_c takes no sss: S Psi is rotated explicitly
        zrayleigh_ritz(zc.z, a.neigmx, vs, l_rrb, psi, spsi, cps, (double2 *)Hp.v(), sv);
        zc.c.q_zero = qz;
      };
      if (a.l_rrb) rr(true);
      if (use_cg) {
        zkscg_c(zc.z, a.neigmx, a.northo, a.nkscg, a.nsdmax, a.nprecon_cg, a.key_cmpt, a.key_nocmpt,
                a.epssd, vs, psi, spsi, cps, (double2 *)Hp.v(), sv, re, a.ksconv, a.ksitmax);
      } else {  // scf_diag8f.F90:241-259
        const ZDiisArgs da{a.neigmx, a.northo, a.ndiismax, a.nprecon_diis, a.key_cmpt, a.l_hsv,
                           a.eps_eig_diis, a.alambda_diis, a.alambda_min, a.alambda_max, a.epssd};
        zksdiis_c(zc.z, da, vs, psi, spsi, cps, (double2 *)Hp.v(), sv, re, a.ksconv, a.ksitmax);
      }
      if (a.l_rra) rr(false);
    }
    if (a.l_reo)
      for (int ns = 1; ns <= a.nums; ns++)
        zreorder(zc.z, a.neigmx, Sv.d() + kofs(1, a.neigmx, a.nums, ns, nk),
                 Re.d() + kofs(1, a.neigmx, a.nums, ns, nk),
                 (double2 *)Ps.v() + kofs(n, a.neigmx, a.nums, ns, nk),
                 (double2 *)Sp.v() + kofs(n, a.neigmx, a.nums, ns, nk),
                 (double2 *)Cs.v() + kofs(n1, a.neigmx, a.nums, ns, nk), (double2 *)Hp.v());
  }
  Cs.finish(); Ps.finish(); Sp.finish(); Hp.finish(); Sv.finish(); Re.finish();
}

// reorderstates with nrc = 1 (reorderstates8f.F90:70-124)
void reorder_bloch(int num_ppcell, int nprjmx, int neigmx, int nums, int ncol, int numk, int ncpx,
                   int ncpy, int ncpz, int ns1, int nk, double *sval, double *residual_states,
                   void *cspsep, void *sveccm, void *ssvcm, void *hsvcm) {
  need_collinear(ncol, 0, "reorderstates");
  Grid g{ncpx, ncpy, ncpz, 1, 3, 1.0, 1.0, 1.0};
  ZCall zc(g, 0.0, 0.0, 0.0);
  zc.c.pl.nprjmx = nprjmx;
  zc.c.pl.num_ppcell = num_ppcell;
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  Arr Sv(sval, (size_t)nall * D, Arr::InOut);
  Arr Re(residual_states, (size_t)nall * D, Arr::InOut);
  Arr Cs(cspsep, (size_t)n1 * nall * Z, Arr::InOut);
  Arr Ps(sveccm, (size_t)n * nall * Z, Arr::InOut);
  Arr Sp(ssvcm, (size_t)n * nall * Z, Arr::InOut);
  Arr Hp(hsvcm, (size_t)n * neigmx * Z, Arr::InOut);
  zreorder(zc.z, neigmx, Sv.d() + kofs(1, neigmx, nums, ns1, nk), Re.d() + kofs(1, neigmx, nums, ns1, nk),
           (double2 *)Ps.v() + kofs(n, neigmx, nums, ns1, nk),
           (double2 *)Sp.v() + kofs(n, neigmx, nums, ns1, nk),
           (double2 *)Cs.v() + kofs(n1, neigmx, nums, ns1, nk), (double2 *)Hp.v());
  Sv.finish(); Re.finish(); Cs.finish(); Ps.finish(); Sp.finish(); Hp.finish();
}

}  // namespace kukan
