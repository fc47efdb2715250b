// abi_trans.hip — extern "C" entry points of potential assembly (declared in
// include/kukan_hip.h): the coarse <-> dense transfers of trans8f.F90 and their building blocks,
// the augmented dense density, the V_eff sum, and the grid chain of scf_potentials from the
// smooth density to V_eff in one call.
#include <algorithm>
#include <vector>

#include "abi_common.hpp"
#include "kukan_hip.h"
#include "transfer.hpp"

using namespace kukan;
using namespace kukan::abi;

namespace {

// the value trans_d2c_vxc puts on the isolated faces of the dense halo (trans8f.F90:429)
constexpr double VXC_FAR = -1.865282921333129e-005;

struct Work {
  const TransferGrid &t;
  double *vc, *vd, *w1, *w2;
  explicit Work(const TransferGrid &tg)
      : t(tg),
        vc((double *)scratch("tr.vc", sizeof(double) * tg.coarse_halo_count())),
        vd((double *)scratch("tr.vd", sizeof(double) * tg.dense_halo_count())),
        w1((double *)scratch("tr.w1", sizeof(double) * tg.dense_halo_count())),
        w2((double *)scratch("tr.w2", sizeof(double) * tg.dense_halo_count())) {}
  // coarse interior -> dense interior (trans_c2d_smtcharge, one component)
  void c2d(const double *in, double *out, hipStream_t s) const {
    t.fill(in, false, vc, s);
    t.prolong(vc, out, false, w1, w2, s);
  }
  // vd holds the halo'd dense field: restrict, fold, interior (one component)
  void d2c(double *out, hipStream_t s) const {
    t.restrict_(vd, vc, w1, w2, s);
    t.fold(vc, s);
    t.coarse_out(vc, out, s);
  }
  void zero_dense(hipStream_t s) const { HIPCK(hipMemsetAsync(vd, 0, sizeof(double) * t.dense_halo_count(), s)); }
};

// the per-atom factor cp(1, indspe(na)) and the multipole tables on the device
struct Multipoles {
  Arr mp, cp1, ax, ay, az;
  Multipoles(int natom, int num_spe, const int *indspe, const double *atmpole, const double *cp,
             const double *atx, const double *aty, const double *atz) {
    const std::vector<int> isp = host_ints(indspe, natom);
    const std::vector<double> cph = host_doubles(cp, (size_t)8 * num_spe);
    cph1.resize(std::max(1, natom));
    for (int na = 0; na < natom; na++) {
      if (isp[na] < 1 || isp[na] > num_spe) throw Error(1, "trans_d2c_vh: indspe out of range");
      cph1[na] = cph[8 * (size_t)(isp[na] - 1)];
    }
    mp = Arr(atmpole, (size_t)10 * natom * D, Arr::In);
    cp1 = Arr(cph1.data(), cph1.size() * D, Arr::In);
    ax = Arr(atx, (size_t)natom * D, Arr::In);
    ay = Arr(aty, (size_t)natom * D, Arr::In);
    az = Arr(atz, (size_t)natom * D, Arr::In);
  }
  std::vector<double> cph1;
};

// scf_rhoaugdense on device arrays
void rhoaugdense(int key_natpri_inps, int key_pp_paw, int natom, int num_spe, int nspv, long nd,
                 int num_list_d, int num_ppcell_d, const int *indspe, const int *natprid,
                 const int *napsd, const int *natinfd, const int *ntyppp, const int *lstvecd2,
                 const double *rho_dense_dev, const double *rhoaug3d, double *out_dev, hipStream_t s) {
  AugPlan pl;
  pl.build(key_natpri_inps, key_pp_paw, natom, nd, num_list_d, host_ints(indspe, natom),
           host_ints(natprid, natom), host_ints(napsd, natom), host_ints(natinfd, natom),
           host_ints(ntyppp, num_spe), host_ints(lstvecd2, (size_t)num_list_d * num_ppcell_d));
  spin_sum(nd, nspv, rho_dense_dev, out_dev, s);
  if (pl.nrows == 0) return;
  Arr aug(rhoaug3d, (size_t)num_list_d * num_ppcell_d * D, Arr::In);
  Arr rows(pl.rows_h.data(), pl.rows_h.size() * I, Arr::In);
  Arr rptr(pl.row_ptr_h.data(), pl.row_ptr_h.size() * I, Arr::In);
  Arr ent(pl.ent_h.data(), pl.ent_h.size() * I, Arr::In);
  aug_add(pl, rows.i(), rptr.i(), ent.i(), aug.d(), out_dev, s);
  sync();  // the plan's host arrays go out of scope
}

}  // namespace

extern "C" {

int kukan_interpolation(int ncpx, int ncpy, int ncpz, int nmesh, double *v_dd, double *v_cc,
                        int nperi, int nf, int is, int ndisp) {
  (void)ndisp;
  API_BEGIN
  if (is != 1 && is != 2) throw Error(1, "error in interpolation! IS should be 1 or 2!");
  TransferGrid t(ncpx, ncpy, ncpz, nf, nmesh, nperi);
  Work w(t);
  Arr vd(v_dd, t.dense_halo_count() * D, Arr::InOut);
  Arr vc(v_cc, t.coarse_halo_count() * D, Arr::InOut);
  if (is == 1) t.prolong(vc.d(), vd.d(), true, w.w1, w.w2, lib().stream);
  else t.restrict_(vd.d(), vc.d(), w.w1, w.w2, lib().stream);
  vd.finish(); vc.finish();
  API_END
}

int kukan_overlap_interpolation(int ndisp, int nperi, int ncpa, int ncpb, int ncpc, int nf, int nf1,
                                double *v_cc, int nflag) {
  (void)ndisp;
  API_BEGIN
  if (nflag != 1 && nflag != 2) throw Error(1, "error in overlap_interpolation. nflag should be 1 or 2!");
  if (nf1 != nf - 1) throw Error(1, "overlap_interpolation: nf1 must be nf - 1");
  TransferGrid t(ncpa, ncpb, ncpc, nf, 1, nperi);
  Arr v(v_cc, t.coarse_halo_count() * D, Arr::InOut);
  if (nflag == 1) {
    double *tmp = (double *)scratch("tr.ovl", sizeof(double) * t.coarse_halo_count());
    HIPCK(hipMemcpyAsync(tmp, v.d(), sizeof(double) * t.coarse_halo_count(), hipMemcpyDeviceToDevice,
                         lib().stream));
    t.fill(tmp, true, v.d(), lib().stream);
  } else {
    t.fold(v.d(), lib().stream);
  }
  v.finish();
  API_END
}

int kukan_trans_c2d_smtcharge(int nmesh, int nspv, int nperi, int ndisp, int nf, int ncpx, int ncpy,
                              int ncpz, const double *rhosmt, double *rho_dense) {
  (void)ndisp;
  API_BEGIN
  TransferGrid t(ncpx, ncpy, ncpz, nf, nmesh, nperi);
  Work w(t);
  const long n = t.coarse_count(), nd = t.dense_count();
  Arr in(rhosmt, (size_t)n * nspv * D, Arr::In);
  Arr out(rho_dense, (size_t)nd * nspv * D, Arr::Out);
  for (int ns = 0; ns < nspv; ns++) w.c2d(in.d() + n * ns, out.d() + nd * ns, lib().stream);
  out.finish();
  API_END
}

int kukan_trans_d2c_smtcharge(int nmesh, int nperi, int ndisp, int nf, int ncpx, int ncpy, int ncpz,
                              const double *rho_aug_dense, double *rho_coarse) {
  (void)ndisp;
  API_BEGIN
  TransferGrid t(ncpx, ncpy, ncpz, nf, nmesh, nperi);
  Work w(t);
  Arr in(rho_aug_dense, t.dense_count() * D, Arr::In);
  Arr out(rho_coarse, t.coarse_count() * D, Arr::Out);
  w.zero_dense(lib().stream);
  t.dense_in(in.d(), w.vd, lib().stream);
  w.d2c(out.d(), lib().stream);
  out.finish();
  API_END
}

int kukan_trans_d2c_vh(int natom, int num_spe, int nmesh, int nperi, int ndisp, int nf, int ncpx,
                       int ncpy, int ncpz, double xmax, double ymax, double zmax, const int *indspe,
                       const double *atmpole, const double *cp, const double *atx, const double *aty,
                       const double *atz, double *vh_coarse, const double *vh_dense) {
  (void)ndisp;
  API_BEGIN
  TransferGrid t(ncpx, ncpy, ncpz, nf, nmesh, nperi);
  Work w(t);
  Arr in(vh_dense, t.dense_count() * D, Arr::In);
  Arr out(vh_coarse, t.coarse_count() * D, Arr::Out);
  w.zero_dense(lib().stream);
  t.dense_in(in.d(), w.vd, lib().stream);
  if (nperi < 3) {
    Multipoles m(natom, num_spe, indspe, atmpole, cp, atx, aty, atz);
    t.multipole_halo(natom, xmax, ymax, zmax, m.mp.d(), m.cp1.d(), m.ax.d(), m.ay.d(), m.az.d(), w.vd,
                     lib().stream);
    sync();  // m's host table goes out of scope
  }
  w.d2c(out.d(), lib().stream);
  out.finish();
  API_END
}

int kukan_trans_d2c_vxc(int nmesh, int nspv, int nperi, int ndisp, int nf, int ncpx, int ncpy,
                        int ncpz, double *vx, const double *vx_dense) {
  (void)ndisp;
  API_BEGIN
  TransferGrid t(ncpx, ncpy, ncpz, nf, nmesh, nperi);
  Work w(t);
  const long n = t.coarse_count(), nd = t.dense_count();
  Arr in(vx_dense, (size_t)nd * nspv * D, Arr::In);
  Arr out(vx, (size_t)n * nspv * D, Arr::Out);
  w.zero_dense(lib().stream);
  t.const_halo(VXC_FAR, w.vd, lib().stream);
  for (int ns = 0; ns < nspv; ns++) {
    t.dense_in(in.d() + nd * ns, w.vd, lib().stream);
    w.d2c(out.d() + n * ns, lib().stream);
  }
  out.finish();
  API_END
}

int kukan_scf_rhoaugdense(int key_natpri_inps, int key_pp_paw, int natom, int num_spe, int nspv,
                          int ncpx_d, int ncpy_d, int ncpz_d, int num_list_d, int num_ppcell_d,
                          const int *indspe, const int *natprid, const int *napsd,
                          const int *natinfd, const int *ntyppp, const int *lstvecd2,
                          const double *rho_dense, const double *rhoaug3d, double *rho_aug_dense) {
  API_BEGIN
  const long nd = (long)ncpx_d * ncpy_d * ncpz_d;
  Arr in(rho_dense, (size_t)nd * nspv * D, Arr::In);
  Arr out(rho_aug_dense, (size_t)nd * D, Arr::Out);
  rhoaugdense(key_natpri_inps, key_pp_paw, natom, num_spe, nspv, nd, num_list_d, num_ppcell_d, indspe,
              natprid, napsd, natinfd, ntyppp, lstvecd2, in.d(), rhoaug3d, out.d(), lib().stream);
  out.finish();
  API_END
}

int kukan_scf_veff(int nspv, int ncpx, int ncpy, int ncpz, double xmax, double ymax, double zmax,
                   double dx, double dy, double dz, double biasx, double biasy, double biasz,
                   const double *vloc_coarse, const double *vh_coarse, const double *vx,
                   double *veff) {
  API_BEGIN
  const long n = (long)ncpx * ncpy * ncpz;
  Arr vl(vloc_coarse, (size_t)n * D, Arr::In), vh(vh_coarse, (size_t)n * D, Arr::In);
  Arr x(vx, (size_t)n * nspv * D, Arr::In);
  Arr out(veff, (size_t)n * nspv * D, Arr::Out);
  veff_sum(ncpx, ncpy, ncpz, nspv, xmax, ymax, zmax, dx, dy, dz, biasx, biasy, biasz, vl.d(), vh.d(), x.d(),
           out.d(), lib().stream);
  out.finish();
  API_END
}

int kukan_scf_potentials_smooth(int key_natpri_inps, int key_pp_paw, int nmesh, int ncpx, int ncpy,
                                int ncpz, int natom, int num_spe, int num_list_d, int num_ppcell_d,
                                int nspv, int ndisp, int nperi, int nf, int nfh, int ncgres,
                                int ncgmin, int ncgmax, const int *indspe, const int *natprid,
                                const int *napsd, const int *natinfd, const int *ntyppp,
                                const int *lstvecd2, double epsvh, double xmax, double ymax,
                                double zmax, double dx, double dy, double dz, double biasx,
                                double biasy, double biasz, const double *cp, const double *atx,
                                const double *aty, const double *atz, const double *atmpole,
                                const double *rhosmt, const double *rhoaug3d,
                                const double *vloc_coarse, const double *vx, const double *boundx,
                                const double *boundy, const double *boundz, double *rho_aug_dense,
                                double *vh_dense, double *vh_coarse, double *veff) {
  (void)ndisp;
  API_BEGIN
  if (nfh != 4 && nfh != 5 && nfh != 6)
    throw Error(6, "error in scf_hartree! This subroutine is for nfh ne 4, 5 or 6.");
  TransferGrid t(ncpx, ncpy, ncpz, nf, nmesh, nperi);
  Work w(t);
  hipStream_t s = lib().stream;
  const long n = t.coarse_count(), nd = t.dense_count();
  const int nxd = ncpx * nmesh, nyd = ncpy * nmesh, nzd = ncpz * nmesh;
  const int nd_flag = nperi / 3;
  Arr rho(rhosmt, (size_t)n * nspv * D, Arr::In);
  Arr vl(vloc_coarse, (size_t)n * D, Arr::In);
  Arr x(vx, (size_t)n * nspv * D, Arr::In);
  const size_t bx = nd_flag ? 1 : (size_t)2 * nfh * nyd * nzd;
  const size_t by = nd_flag ? 1 : (size_t)nxd * 2 * nfh * nzd;
  const size_t bz = nd_flag ? 1 : (size_t)nxd * nyd * 2 * nfh;
  Arr Bx(boundx, bx * D, Arr::In), By(boundy, by * D, Arr::In), Bz(boundz, bz * D, Arr::In);
  Arr rad(rho_aug_dense, (size_t)nd * D, Arr::Out);
  Arr vhd(vh_dense, (size_t)nd * D, Arr::InOut);
  Arr vhc(vh_coarse, (size_t)n * D, Arr::Out);
  Arr ve(veff, (size_t)n * nspv * D, Arr::Out);
  // the smooth density on the dense grid (scf_potentials8f.F90:115-118)
  double *rho_dense = (double *)scratch("tr.rhod", sizeof(double) * nd * std::max(1, nspv));
  {
    Timed tm("pot_c2d", 0.0);
    for (int ns = 0; ns < nspv; ns++) w.c2d(rho.d() + n * ns, rho_dense + nd * ns, s);
  }
  // plus the augmentation charge (:130-136)
  {
    Timed tm("pot_rhoaug", 0.0);
    rhoaugdense(key_natpri_inps, key_pp_paw, natom, num_spe, nspv, nd, num_list_d, num_ppcell_d, indspe,
                natprid, napsd, natinfd, ntyppp, lstvecd2, rho_dense, rhoaug3d, rad.d(), s);
  }
  // Hartree on the dense grid, zero average for the periodic cell (:159-168)
  {
    Timed tm("pot_hartree", 0.0);
    HartreeArgs a{nd_flag, nxd, nyd, nzd, nfh, nperi, ncgres, ncgmin, ncgmax, xmax, ymax, zmax, epsvh,
                  rad.d(), Bx.d(), By.d(), Bz.d(), vhd.d()};
    const int loops = hartree(a);
    lib().params["hartree_iterations"] = loops;
    if (nperi == 3) {
      const double ddx = dx / nmesh, ddy = dy / nmesh, ddz = dz / nmesh;
      correct_potential(nd, ddx * ddy * ddz / (8.0 * xmax * ymax * zmax), vhd.d());
    }
  }
  // back to the coarse grid (:170-178) and the V_eff sum (:255-289)
  {
    Timed tm("pot_d2c", 0.0);
    w.zero_dense(s);
    t.dense_in(vhd.d(), w.vd, s);
    if (nperi < 3) {
      Multipoles m(natom, num_spe, indspe, atmpole, cp, atx, aty, atz);
      t.multipole_halo(natom, xmax, ymax, zmax, m.mp.d(), m.cp1.d(), m.ax.d(), m.ay.d(), m.az.d(), w.vd, s);
      sync();
    }
    w.d2c(vhc.d(), s);
    veff_sum(ncpx, ncpy, ncpz, nspv, xmax, ymax, zmax, dx, dy, dz, biasx, biasy, biasz, vl.d(), vhc.d(),
             x.d(), ve.d(), s);
  }
  rad.finish(); vhd.finish(); vhc.finish(); ve.finish();
  API_END
}

}  // extern "C"
