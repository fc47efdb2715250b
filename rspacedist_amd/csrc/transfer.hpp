// transfer.hpp — the coarse <-> dense grid transfers of potential assembly (transfer.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace kukan {

// One coarse grid (nx, ny, nz) with its dense grid (nmesh per axis), the interpolation order
// 2 nf and the periodicity nperi.  Halo'd coarse arrays span -(nf-1) .. n+nf per axis, halo'd
// dense arrays -(nf nmesh - 1) .. n nmesh + nf nmesh, both column-major.
class TransferGrid {
 public:
  TransferGrid(int nx, int ny, int nz, int nf, int nmesh, int nperi);
  ~TransferGrid();
  TransferGrid(const TransferGrid &) = delete;
  TransferGrid &operator=(const TransferGrid &) = delete;
  long coarse_halo_count() const;
  long dense_halo_count() const;
  long dense_count() const;
  long coarse_count() const;
  // work sizes (doubles) of prolong / restrict_
  long pro_t1() const { return dense_halo_count(); }
  long pro_t2() const { return dense_halo_count(); }
  // the one-rank halo fill (overlap_interpolation, nflag = 1) of an interior (in_halo = false,
  // isolated halos zero) or halo'd array into out_halo
  void fill(const double *in, bool in_halo, double *out_halo, hipStream_t s) const;
  // overlap_interpolation, nflag = 2, in place
  void fold(double *vc_halo, hipStream_t s) const;
  // interpolation, is = 1: dense interior (out_halo: the whole halo'd dense array, zeroed first)
  void prolong(const double *vc_halo, double *out, bool out_halo, double *t1, double *t2,
               hipStream_t s) const;
  // interpolation, is = 2 (clears vd_halo's periodic halo first, as the reference does)
  void restrict_(double *vd_halo, double *vc_halo, double *t2r, double *t1r, hipStream_t s) const;
  void dense_in(const double *src, double *vd_halo, hipStream_t s) const;
  void coarse_out(const double *vc_halo, double *dst, hipStream_t s) const;
  // trans_d2c_vh's multipole halo; cp1[na] = cp(1, indspe(na))
  void multipole_halo(int natom, double xmax, double ymax, double zmax, const double *atmpole,
                      const double *cp1, const double *atx, const double *aty, const double *atz,
                      double *vd_halo, hipStream_t s) const;
  // trans_d2c_vxc's constant on the isolated faces
  void const_halo(double c, double *vd_halo, hipStream_t s) const;

 private:
  struct Impl;
  Impl *impl;
};

// augmentation entries per dense point, in the reference's atom / list order
struct AugPlan {
  long nrows = 0;
  std::vector<int> rows_h, row_ptr_h, ent_h;
  void build(int key_natpri_inps, int key_pp_paw, int natom, long nd, int num_list_d,
             const std::vector<int> &indspe, const std::vector<int> &natprid,
             const std::vector<int> &napsd, const std::vector<int> &natinfd,
             const std::vector<int> &ntyppp, const std::vector<int> &lstvecd2);
};
void spin_sum(long n, int nspv, const double *rho, double *out, hipStream_t s);
void aug_add(const AugPlan &pl, const int *rows, const int *row_ptr, const int *ent, const double *aug,
             double *out, hipStream_t s);
void veff_sum(int nx, int ny, int nz, int nspv, double xmax, double ymax, double zmax, double dx,
              double dy, double dz, double bx, double by, double bz, const double *vloc,
              const double *vh, const double *vx, double *veff, hipStream_t s);

}  // namespace kukan
