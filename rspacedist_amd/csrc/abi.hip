// abi.hip — extern "C" entry points of libkukan_hip.so (declared in include/kukan_hip.h).
//
// Every entry point keeps the argument list of the reference routine it replaces; arrays
// are resolved to device memory (in place for device pointers and bound arrays, staged
// copies otherwise), the device implementation in solvers.hip runs on the library stream,
// and staged outputs are copied back before return.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "kukan_hip.h"
#include "abi_common.hpp"
#include "kukan_internal.hpp"
#include "solvers.hpp"
#include "bloch.hpp"

using namespace kukan;

using namespace kukan::abi;

extern "C" {

int kukan_init(int device) {
  try {
    if (lib().stream && device != lib().device) throw Error(1, "kukan_init: already initialised on another device");
    lib().device = device;
    ensure_init();
    noalloc_setup(lib().device);
    return 0;
  } catch (const std::exception &e) {
    lib().last_error = e.what();
    return 1;
  }
}

int kukan_finalize(void) {
  Lib &L = lib();
  if (!L.stream) return 0;
  (void)hipStreamSynchronize(L.stream);
  for (auto &kv : L.scratch) (void)hipFree(kv.second.first);
  L.scratch.clear();
  for (auto &kv : L.bound) (void)hipFree(kv.second.dev);
  L.bound.clear();
  rocblas_destroy_handle(L.blas);
  (void)hipStreamDestroy(L.stream);
  L.stream = nullptr;
  L.blas = nullptr;
  return 0;
}

const char *kukan_last_error(void) { return lib().last_error.c_str(); }

int kukan_set_param(const char *key, double value) {
  lib().params[key] = value;
  return 0;
}
double kukan_get_param(const char *key) { return param(key, 0.0); }

void *kukan_stream(void) {
  try {
    ensure_init();
  } catch (...) {
    return nullptr;
  }
  return (void *)lib().stream;
}

int kukan_bind(void *host_ptr, size_t bytes) {
  API_BEGIN
  if (lib().bound.count(host_ptr)) throw Error(1, "kukan_bind: array already bound");
  void *d = nullptr;
  HIPCK(hipMalloc(&d, bytes));
  HIPCK(hipMemcpyAsync(d, host_ptr, bytes, hipMemcpyHostToDevice, lib().stream));
  lib().bound[host_ptr] = {d, bytes};
  API_END
}
int kukan_unbind(void *host_ptr) {
  API_BEGIN
  auto it = lib().bound.find(host_ptr);
  if (it == lib().bound.end()) throw Error(1, "kukan_unbind: array not bound");
  HIPCK(hipMemcpyAsync(host_ptr, it->second.dev, it->second.bytes, hipMemcpyDeviceToHost, lib().stream));
  sync();
  HIPCK(hipFree(it->second.dev));
  lib().bound.erase(it);
  API_END
}
int kukan_sync_to_host(void *host_ptr) {
  API_BEGIN
  auto it = lib().bound.find(host_ptr);
  if (it == lib().bound.end()) throw Error(1, "kukan_sync_to_host: array not bound");
  HIPCK(hipMemcpyAsync(host_ptr, it->second.dev, it->second.bytes, hipMemcpyDeviceToHost, lib().stream));
  API_END
}
int kukan_sync_to_device(void *host_ptr) {
  API_BEGIN
  auto it = lib().bound.find(host_ptr);
  if (it == lib().bound.end()) throw Error(1, "kukan_sync_to_device: array not bound");
  HIPCK(hipMemcpyAsync(it->second.dev, host_ptr, it->second.bytes, hipMemcpyHostToDevice, lib().stream));
  API_END
}

int kukan_device_malloc(void **ptr, size_t bytes) {
  API_BEGIN
  *ptr = nullptr;
  HIPCK(hipMalloc(ptr, bytes ? bytes : 1));
  API_END
}
int kukan_device_free(void *ptr) {
  API_BEGIN
  sync();
  HIPCK(hipFree(ptr));
  API_END
}
int kukan_memcpy(void *dst, const void *src, size_t bytes, int kind) {
  API_BEGIN
  const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                          : kind == 2 ? hipMemcpyDeviceToHost
                                      : hipMemcpyDeviceToDevice;
  if (kind < 1 || kind > 3) throw Error(1, "kukan_memcpy: kind must be 1, 2 or 3");
  if (kind == 3 && bytes % 8 == 0 && ((uintptr_t)dst | (uintptr_t)src) % 8 == 0 &&
      is_device_ptr(dst) && is_device_ptr(src)) {
    // device -> device through the library's own copy kernel (a known-traffic kernel:
    // 8 bytes read and 8 written per element, used to calibrate HBM counters)
    const long n = (long)(bytes / 8);
    Timed t("copy", 16.0 * (double)n);
    launch_copy2d(n, 1, (const double *)src, n, (double *)dst, n, lib().stream);
  } else {
    HIPCK(hipMemcpyAsync(dst, src, bytes, k, lib().stream));
  }
  API_END
}
int kukan_fill_normal(double *dev, size_t n, uint64_t seed) {
  API_BEGIN
  launch_fill_normal(dev, (long)n, (unsigned long long)seed, lib().stream);
  API_END
}
int kukan_device_synchronize(void) {
  API_BEGIN
  API_END
}

int kukan_subspace_gemm(int trans_a, int m, int n, long k, double alpha, const double *A,
                        long lda, const double *B, long ldb, double beta, double *C, long ldc) {
  API_BEGIN
  if (m < 0 || n < 0 || k < 0) throw Error(1, "subspace_gemm: negative dimension");
  if (m == 0 || n == 0) return 0;
  const long arows = trans_a ? k : m, acols = trans_a ? m : k;
  if (lda < std::max<long>(1, arows) || ldb < std::max<long>(1, k) || ldc < m)
    throw Error(1, "subspace_gemm: leading dimension too small");
  Arr a(A, sizeof(double) * (size_t)(lda * (acols - 1) + arows), Arr::In);
  Arr b(B, sizeof(double) * (size_t)(ldb * (n - 1) + k), Arr::In);
  Arr c(C, sizeof(double) * (size_t)(ldc * (n - 1) + m), Arr::InOut);
  if (!trans_a && c.d() == a.d()) throw Error(1, "subspace_gemm: C must not alias A");
  if (k == 0) {
    subspace_tn(m, n, 0, 0.0, a.d(), lda, b.d(), ldb, beta, c.d(), ldc);
  } else if (trans_a) {
    subspace_tn(m, n, k, alpha, a.d(), lda, b.d(), ldb, beta, c.d(), ldc);
  } else {
    subspace_nn(m, (int)k, n, alpha, a.d(), lda, b.d(), ldb, beta, c.d(), ldc);
  }
  c.finish();
  API_END
}

int kukan_timing_get(const char *tag, double *ms, long *count, double *work) {
  API_BEGIN
  timing_collect();
  auto &t = timing_table();
  auto it = t.find(tag);
  *ms = it == t.end() ? 0.0 : it->second.ms;
  *count = it == t.end() ? 0 : it->second.count;
  *work = it == t.end() ? 0.0 : it->second.work;
  API_END
}
int kukan_timing_reset(void) {
  API_BEGIN
  timing_collect();
  timing_table().clear();
  API_END
}
// Tags with recorded time, separated by ';' (buffer of len bytes).
int kukan_timing_tags(char *buf, size_t len) {
  API_BEGIN
  timing_collect();
  std::string s;
  for (auto &kv : timing_table()) s += kv.first + ";";
  if (len) {
    strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  API_END
}

// ----------------------------------------------------------------------------- leaves
int kukan_tools_definefdcoef(int nf, double *acfd) {
  double a[9];
  fd_coefficients(nf, a);
  if (is_device_ptr(acfd)) {
    API_BEGIN
    HIPCK(hipMemcpyAsync(acfd, a, sizeof a, hipMemcpyHostToDevice, lib().stream));
    API_END
  }
  memcpy(acfd, a, sizeof a);
  return 0;
}

int kukan_overlap_finitedifference_r(int nperi, int ncpa, int ncpb, int ncpc, int nf, int nf1,
                                     int nsnd, double *v_cc) {
  API_BEGIN
  const size_t cnt = (size_t)(ncpa + nf + nf1 + 1) * (ncpb + nf + nf1 + 1) * (ncpc + nf + nf1 + 1);
  Arr v(v_cc, cnt * D, Arr::InOut);
  launch_halo_fill(v.d(), nperi, ncpa, ncpb, ncpc, nf, nf1, nsnd, lib().stream);
  v.finish();
  API_END
}

int kukan_overlap_fdcheck_r(int nperi, int ncpa, int ncpb, int ncpc, int nf, int nf1, int nsnd,
                            double *v_cc) {
  (void)nperi; (void)ncpa; (void)ncpb; (void)ncpc; (void)nf; (void)nf1; (void)nsnd; (void)v_cc;
  return 0;  // the fill completed in kukan_overlap_finitedifference_r
}

int kukan_kslaplacian_r(int ncpx, int ncpy, int ncpz, int neigmx, int nf, double dx, double dy,
                        double dz, const double *vre, const double *veff, int l, double *avre) {
  API_BEGIN
  if (nf < 1 || nf > 8) throw Error(1, "error in kslaplacian_r! This subroutine is for nf in 1-8.");
  Grid g{ncpx, ncpy, ncpz, nf, 3, dx, dy, dz};
  const long n = g.n();
  Arr in(vre, (size_t)(ncpx + 2 * nf) * (ncpy + 2 * nf) * (ncpz + 2 * nf) * D, Arr::In);
  Arr ve(veff, n * D, Arr::In);
  Arr out(avre, (size_t)n * neigmx * D, Arr::InOut);
  StencilCall c;
  c.g = &g; c.in = in.d(); c.ld_in = 0; c.halo_in = true; c.out = out.d() + (size_t)(l - 1) * n;
  c.ld_out = n; c.veff = ve.d(); c.nb = 1; c.accumulate = true;
  launch_stencil(c, lib().stream);
  out.finish();
  API_END
}

int kukan_ksprecondition(int ncpx, int ncpy, int ncpz, int nf, int ncol, const double *vre,
                         double *avre) {
  API_BEGIN
  Grid g{ncpx, ncpy, ncpz, nf, 3, 1.0, 1.0, 1.0};
  const long nh = (long)(ncpx + 2 * nf) * (ncpy + 2 * nf) * (ncpz + 2 * nf), n = g.n();
  Arr in(vre, (size_t)nh * ncol * D, Arr::In);
  Arr out(avre, (size_t)n * ncol * D, Arr::Out);
  launch_precond(g, in.d(), nh, true, nf, out.d(), n, ncol, lib().stream);
  out.finish();
  API_END
}

int kukan_nonlocaloperation_r_01(int natom, int nprjmx, int num_spe, int num_ppcell, int num_list,
                                 int ncpx, int ncpy, int ncpz, int key_natpri_in,
                                 int key_natpri_inps, const int *nprj, const int *indspe,
                                 const int *natinf, const int *natpri, const int *naps,
                                 const int *lstvec2, const double *vnlocp, const double *vre2,
                                 double *rpsep) {
  API_BEGIN
  Grid g{ncpx, ncpy, ncpz, 1, 3, 1.0, 1.0, 1.0};
  Ctx c(g);
  Arr vnl(vnlocp, (size_t)num_list * nprjmx * num_ppcell * D, Arr::In);
  Arr lst(lstvec2, (size_t)num_list * num_ppcell * I, Arr::In);
  Arr v(vre2, (size_t)g.n() * D, Arr::In);
  Arr rp(rpsep, (size_t)nprjmx * num_ppcell * D, Arr::InOut);
  build_plan(c.pl, g, natom, nprjmx, num_spe, num_ppcell, num_list, 1, key_natpri_in,
             key_natpri_inps, nprj, indspe, natinf, natpri, naps, vnl.d(), lst.i(), "plan");
  launch_gather(c.pl, v.d(), g.n(), 1, rp.d(), (long)nprjmx * num_ppcell, true, lib().stream);
  rp.finish();
  API_END
}

int kukan_nonlocaloperation_r_02(int nhs, int natom, int num_spe, int nums, int nprjmx, int ns,
                                 int num_list, int num_ppcell, int ncpx, int ncpy, int ncpz,
                                 int key_natpri_in, int key_natpri_inps, double dx, double dy,
                                 double dz, const int *nprj, const int *indspe, const int *natinf,
                                 const int *lstvec2, const int *natpri, const int *naps,
                                 const double *vnlocp, const double *rpsep, const double *dijsss,
                                 double *avre, double *avr) {
  (void)avr;
  API_BEGIN
  if (nhs != 1 && nhs != 0) throw Error(1, "error occurs in nonlocaloperation_r_02");
  if (nhs == 0 && ns != 1) throw Error(1, "nonlocaloperation_r_02: nhs=0 and ns/=1");
  Grid g{ncpx, ncpy, ncpz, 1, 3, dx, dy, dz};
  Ctx c(g);
  const long n1 = (long)nprjmx * num_ppcell;
  const size_t xcnt = (size_t)nprjmx * nprjmx * (nums * nhs - nhs + 1) * (natom * nhs + num_spe * (1 - nhs));
  Arr vnl(vnlocp, (size_t)num_list * nprjmx * num_ppcell * D, Arr::In);
  Arr lst(lstvec2, (size_t)num_list * num_ppcell * I, Arr::In);
  Arr rp(rpsep, (size_t)n1 * D, Arr::In);
  Arr X(dijsss, xcnt * D, Arr::In);
  Arr out(avre, (size_t)g.n() * D, Arr::InOut);
  build_plan(c.pl, g, natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in,
             key_natpri_inps, nprj, indspe, natinf, natpri, naps, vnl.d(), lst.i(), "plan");
  double *coef = c.buf("coef", (size_t)std::max<long>(1, n1));
  launch_coef(c.pl, rp.d(), n1, 1, X.d(), nhs == 0, ns, 1.0 / (dx * dy * dz), coef, n1, lib().stream);
  launch_expand(c.pl, g.n(), coef, n1, 1, out.d(), g.n(), lib().stream);
  out.finish();
  API_END
}

// ----------------------------------------------------------------------------- coarse
int kukan_kscg_r(int natom, int num_spe, int neigmx, int nums, int nspv, int numk, int nperi,
                 int northo, int ns, int nk, int nf, int nkscg, int nsdmax, int nprecon_cg,
                 int nprjmx, int num_list, int num_ppcell, int ndisp, int ncpx, int ncpy, int ncpz,
                 int key_ortho_nocmpt_innerproduct, int key_ortho_cmpt_innerproduct,
                 int key_natpri_in, int key_natpri_inps, int key_natpri_out, double dx, double dy,
                 double dz, double epssd, const int *indspe, const int *natpri, const int *naps,
                 const int *nprj, const int *natinf, const int *lstvec2, const int *latom,
                 const int *ntyppp, const double *dij, const double *veff, const double *vnlocp,
                 const double *sss, int *ksconv, int *ksitmax, double *sval, double *rspsep,
                 double *svecre, double *ssvre, double *residual_states, double *hsvre) {
  (void)ndisp; (void)key_natpri_out; (void)latom; (void)ntyppp;
  API_BEGIN
  Grid g{ncpx, ncpy, ncpz, nf, nperi, dx, dy, dz};
  Ctx c(g);
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  Arr vnl(vnlocp, (size_t)num_list * nprjmx * num_ppcell * D, Arr::In);
  Arr lst(lstvec2, (size_t)num_list * num_ppcell * I, Arr::In);
  Arr Dij(dij, (size_t)nprjmx * nprjmx * nums * natom * D, Arr::In);
  Arr Sss(sss, (size_t)nprjmx * nprjmx * num_spe * D, Arr::In);
  Arr Ve(veff, (size_t)n * nspv * D, Arr::In);
  Arr Sv(sval, (size_t)nall * D, Arr::InOut);
  Arr Rs(rspsep, (size_t)n1 * nall * D, Arr::InOut);
  Arr Ps(svecre, (size_t)n * nall * D, Arr::InOut);
  Arr Sp(ssvre, (size_t)n * nall * D, Arr::InOut);
  Arr Re(residual_states, (size_t)neigmx * D, Arr::Out);
  Arr Hp(hsvre, (size_t)n * neigmx * D, Arr::Out);
  setup(c, natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps,
        nprj, indspe, natinf, natpri, naps, vnl, lst, Dij.d(), Sss.d(), sss);
  const size_t off = (size_t)n * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
  const size_t off1 = (size_t)n1 * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
  const size_t offs = (size_t)neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
  const int nsv = std::min(ns, nspv);
  KscgArgs a{neigmx, ns, northo, nkscg, nsdmax, nprecon_cg, key_ortho_cmpt_innerproduct,
             key_ortho_nocmpt_innerproduct, epssd, Ve.d() + (size_t)n * (nsv - 1),
             Ps.d() + off, Sp.d() + off, Hp.d(), Rs.d() + off1, Sv.d() + offs, Re.d(), ksconv, ksitmax};
  if (param("kscg_mode", 0.0) >= 1.0) kscg_block(c, a, true);
  else kscg_sequential(c, a);
  Sv.finish(); Rs.finish(); Ps.finish(); Sp.finish(); Re.finish(); Hp.finish();
  API_END
}

int kukan_ksdiis_r(int l_hsv, int natom, int num_spe, int neigmx, int nums, int nspv, int numk,
                   int nperi, int ndiismax, int nprecon_diis, int nk, int ns, int nf, int nprjmx,
                   int num_list, int num_ppcell, int ndisp, int northo, int ncpx, int ncpy,
                   int ncpz, int key_ortho_cmpt_innerproduct, int key_natpri_in,
                   int key_natpri_inps, double eps_eig_diis, double alambda_diis,
                   double ratio_diis, double alambda_min, double alambda_max, double epssd,
                   double xmax, double ymax, double zmax, const int *indspe, const int *natpri,
                   const int *naps, const int *nprj, const int *natinf, const int *lstvec2,
                   const int *latom, const int *ntyppp, const double *dij, const double *veff,
                   const double *vnlocp, const double *sss, int *ksconv, int *ksitmax,
                   double *rspsep, double *sval, double *svecre, double *ssvre,
                   double *residual_states, double *hsvre) {
  (void)ndisp; (void)ratio_diis; (void)latom; (void)ntyppp;
  if (nspv != nums) return 0;  // ksdiis8f.F90:78-81
  API_BEGIN
  Grid g{ncpx, ncpy, ncpz, nf, nperi, 2.0 * xmax / ncpx, 2.0 * ymax / ncpy, 2.0 * zmax / ncpz};
  Ctx c(g);
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  Arr vnl(vnlocp, (size_t)num_list * nprjmx * num_ppcell * D, Arr::In);
  Arr lst(lstvec2, (size_t)num_list * num_ppcell * I, Arr::In);
  Arr Dij(dij, (size_t)nprjmx * nprjmx * nums * natom * D, Arr::In);
  Arr Sss(sss, (size_t)nprjmx * nprjmx * num_spe * D, Arr::In);
  Arr Ve(veff, (size_t)n * nspv * D, Arr::In);
  Arr Sv(sval, (size_t)nall * D, Arr::InOut);
  Arr Rs(rspsep, (size_t)n1 * nall * D, Arr::InOut);
  Arr Ps(svecre, (size_t)n * nall * D, Arr::InOut);
  Arr Sp(ssvre, (size_t)n * nall * D, Arr::InOut);
  Arr Re(residual_states, (size_t)neigmx * D, Arr::InOut);
  Arr Hp(hsvre, (size_t)n * neigmx * D, Arr::Out);
  setup(c, natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps,
        nprj, indspe, natinf, natpri, naps, vnl, lst, Dij.d(), Sss.d(), sss);
  const size_t off = (size_t)n * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
  const size_t off1 = (size_t)n1 * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
  const size_t offs = (size_t)neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
  DiisArgs a{neigmx, ns, ndiismax, nprecon_diis, northo, key_ortho_cmpt_innerproduct, l_hsv,
             eps_eig_diis, alambda_diis, alambda_min, alambda_max, epssd,
             Ve.d() + (size_t)n * (ns - 1), Ps.d() + off, Sp.d() + off, Hp.d(), Rs.d() + off1,
             Sv.d() + offs, Re.d(), ksconv, ksitmax};
  ksdiis(c, a);
  Sv.finish(); Rs.finish(); Ps.finish(); Sp.finish(); Re.finish(); Hp.finish();
  API_END
}

int kukan_orthogonalization_r_01(int natom, int neigmx, int nprjmx, int nums, int numk, int nk,
                                 int ns, int l_inp, int num_list, int num_ppcell, int num_spe,
                                 int nflag, int ndisp, int ncpx, int ncpy, int ncpz,
                                 int key_ortho_cmpt_innerproduct, int key_natpri_in,
                                 int key_natpri_inps, double dx, double dy, double dz,
                                 const int *nprj, const int *indspe, const int *natpri,
                                 const int *naps, const int *natinf, const int *lstvec2,
                                 const int *latom, const double *vnlocp, const double *sss,
                                 double *svecre, double *ssvre, double *rspsep) {
  (void)ndisp; (void)latom;
  API_BEGIN
  Grid g{ncpx, ncpy, ncpz, 1, 3, dx, dy, dz};
  Ctx c(g);
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  Arr vnl(vnlocp, (size_t)num_list * nprjmx * num_ppcell * D, Arr::In);
  Arr lst(lstvec2, (size_t)num_list * num_ppcell * I, Arr::In);
  Arr Sss(sss, (size_t)nprjmx * nprjmx * num_spe * D, Arr::In);
  Arr Ps(svecre, (size_t)n * nall * D, Arr::InOut);
  Arr Sp(ssvre, (size_t)n * nall * D, Arr::InOut);
  Arr Rs(rspsep, (size_t)n1 * nall * D, Arr::InOut);
  setup(c, natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps,
        nprj, indspe, natinf, natpri, naps, vnl, lst, nullptr, Sss.d(), sss);
  const size_t off = (size_t)n * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
  const size_t off1 = (size_t)n1 * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
  ortho_r_01(c, neigmx, l_inp, nflag, key_ortho_cmpt_innerproduct, Ps.d() + off, Sp.d() + off,
             Rs.d() + off1);
  Ps.finish(); Sp.finish(); Rs.finish();
  API_END
}

int kukan_orthogonalization_r(int natom, int neigmx, int nprjmx, int nums, int numk,
                              int num_ppcell, int num_spe, int northo, int num_list, int nflag,
                              int ndisp, int ncpx, int ncpy, int ncpz,
                              int key_ortho_cmpt_innerproduct, int key_natpri_in,
                              int key_natpri_inps, double dx, double dy, double dz,
                              const int *nprj, const int *indspe, const int *natpri,
                              const int *naps, const int *natinf, const int *lstvec2,
                              const int *latom, const int *ntyppp, const double *vnlocp,
                              const double *sss, double *svecre, double *ssvre, double *rspsep) {
  (void)ndisp; (void)latom; (void)ntyppp;
  API_BEGIN
  Grid g{ncpx, ncpy, ncpz, 1, 3, dx, dy, dz};
  Ctx c(g);
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  Arr vnl(vnlocp, (size_t)num_list * nprjmx * num_ppcell * D, Arr::In);
  Arr lst(lstvec2, (size_t)num_list * num_ppcell * I, Arr::In);
  Arr Sss(sss, (size_t)nprjmx * nprjmx * num_spe * D, Arr::In);
  Arr Ps(svecre, (size_t)n * nall * D, Arr::InOut);
  Arr Sp(ssvre, (size_t)n * nall * D, Arr::InOut);
  Arr Rs(rspsep, (size_t)n1 * nall * D, Arr::InOut);
  setup(c, natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps,
        nprj, indspe, natinf, natpri, naps, vnl, lst, nullptr, Sss.d(), sss);
  for (int nk = 1; nk <= numk; nk++)
    for (int ns = 1; ns <= nums; ns++) {
      const size_t off = (size_t)n * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
      const size_t off1 = (size_t)n1 * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
      for (int loop = 0; loop < northo; loop++)
        ortho_r_01(c, neigmx, 0, nflag, key_ortho_cmpt_innerproduct, Ps.d() + off, Sp.d() + off,
                   Rs.d() + off1);
    }
  Ps.finish(); Sp.finish(); Rs.finish();
  API_END
}

int kukan_rayleighritz_blacs_ini(int neigmx, int nrc) {
  (void)neigmx; (void)nrc;
  return 0;
}
int kukan_rayleighritz_blacs_fin(void) { return 0; }

int kukan_rayleighritz_r(int natom, int num_ppcell, int num_spe, int nprjmx, int neigmx, int numk,
                         int nums, int nspv, int num_list, int ncpx, int ncpy, int ncpz,
                         int nperi, int nk, int ns, int nf, int ndisp, int key_natpri_in,
                         int key_natpri_inps, double dx, double dy, double dz, const int *nprj,
                         const int *natinf, const int *lstvec2, const int *indspe,
                         const int *natpri, const int *naps, const int *ntyppp, const double *dij,
                         const double *veff, const double *vnlocp, int l_rrb, double *rspsep,
                         double *svecre, double *ssvre, double *hsvre, double *sval) {
  (void)ndisp; (void)ntyppp;
  API_BEGIN
  Grid g{ncpx, ncpy, ncpz, nf, nperi, dx, dy, dz};
  Ctx c(g);
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  Arr vnl(vnlocp, (size_t)num_list * nprjmx * num_ppcell * D, Arr::In);
  Arr lst(lstvec2, (size_t)num_list * num_ppcell * I, Arr::In);
  Arr Dij(dij, (size_t)nprjmx * nprjmx * nums * natom * D, Arr::In);
  Arr Ve(veff, (size_t)n * nspv * D, Arr::In);
  Arr Rs(rspsep, (size_t)n1 * nall * D, Arr::InOut);
  Arr Ps(svecre, (size_t)n * nall * D, Arr::InOut);
  Arr Sp(ssvre, (size_t)n * nall * D, Arr::InOut);
  Arr Hp(hsvre, (size_t)n * neigmx * D, Arr::InOut);
  Arr Sv(sval, (size_t)nall * D, Arr::InOut);
  setup(c, natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps,
        nprj, indspe, natinf, natpri, naps, vnl, lst, Dij.d(), nullptr, nullptr);
  c.q_zero = false;  // the reference rotates S Psi explicitly here (no sss argument)
  const size_t off = (size_t)n * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
  const size_t off1 = (size_t)n1 * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
  const size_t offs = (size_t)neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
  const int nsv = std::min(ns, nspv);
  rayleigh_ritz(c, neigmx, ns, Ve.d() + (size_t)n * (nsv - 1), l_rrb != 0, Ps.d() + off,
                Sp.d() + off, Rs.d() + off1, Hp.d(), Sv.d() + offs, false);
  Rs.finish(); Ps.finish(); Sp.finish(); Hp.finish(); Sv.finish();
  API_END
}

int kukan_reorderstates(int nrc, int num_ppcell, int nprjmx, int neigmx, int nums, int ncol,
                        int numk, int ncpx, int ncpy, int ncpz, int ns1, int nk, double *sval,
                        double *residual_states, double *rspsep, void *cspsep, double *svecre,
                        void *sveccm, double *ssvre, void *ssvcm, double *hsvre, void *hsvcm) {
  API_BEGIN
  if (nrc != 0) {
    reorder_bloch(num_ppcell, nprjmx, neigmx, nums, ncol, numk, ncpx, ncpy, ncpz, ns1, nk, sval,
                  residual_states, cspsep, sveccm, ssvcm, hsvcm);
    sync();
    return 0;
  }
  Grid g{ncpx, ncpy, ncpz, 1, 3, 1.0, 1.0, 1.0};
  Ctx c(g);
  c.pl.nprjmx = nprjmx;
  c.pl.num_ppcell = num_ppcell;
  const int d2 = nums + 1 - ncol;
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  Arr Sv(sval, (size_t)neigmx * d2 * numk * D, Arr::InOut);
  Arr Re(residual_states, (size_t)neigmx * d2 * numk * D, Arr::InOut);
  Arr Rs(rspsep, (size_t)n1 * nall * D, Arr::InOut);
  Arr Ps(svecre, (size_t)n * nall * D, Arr::InOut);
  Arr Sp(ssvre, (size_t)n * nall * D, Arr::InOut);
  Arr Hp(hsvre, (size_t)n * neigmx * D, Arr::InOut);
  const size_t off = (size_t)n * neigmx * ((ns1 - 1) + (size_t)nums * (nk - 1));
  const size_t off1 = (size_t)n1 * neigmx * ((ns1 - 1) + (size_t)nums * (nk - 1));
  const size_t offs = (size_t)neigmx * ((ns1 - 1) + (size_t)d2 * (nk - 1));
  reorder(c, neigmx, Sv.d() + offs, Re.d() + offs, Ps.d() + off, Sp.d() + off, Rs.d() + off1, Hp.d());
  Sv.finish(); Re.finish(); Rs.finish(); Ps.finish(); Sp.finish(); Hp.finish();
  API_END
}

int kukan_scf_diag(int nrc, int nso, int nprjmx, int num_ppcell, int num_spe, int iscf,
                   int ncgscf, int nretcg, int natom, int neigmx, int nums, int ncol, int nspv,
                   int numk, int nperi, int northo, int nf, int ndisp, int num_list, int ncpx,
                   int ncpy, int ncpz, int npxmax, int npymax, int npzmax, int nchange, int nrrz,
                   int key_ortho_nocmpt_innerproduct, int key_ortho_cmpt_innerproduct,
                   int key_natpri_in, int key_natpri_inps, int key_natpri_out, int key_soc_calc,
                   int key_pp_paw, const double *skpxx, const double *skpyy, const double *skpzz,
                   double dx, double dy, double dz, double xmax, double ymax, double zmax,
                   const double *dij, const double *dijsoc, const double *veff,
                   const double *vnlocp, const double *sss, int nkscg, int nsdmax, int nprecon_cg,
                   double epssd, int ndiismax, int nprecon_diis, double eps_eig_diis,
                   double alambda_diis, double ratio_diis, double alambda_min,
                   double alambda_max, const int *indspe, const int *natpri, const int *naps,
                   const int *nprj, const int *natinf, const int *lstvec2, const int *latom,
                   const int *natsoc, const int *ntyppp, const int *lstx, const int *lsty,
                   const int *lstz, const int *natx, const int *naty, const int *natz,
                   double *rspsep, void *cspsep, double *svecre, void *sveccm, double *hsvre,
                   void *hsvcm, double *ssvre, void *ssvcm, double *sval, double *svalsoc,
                   void *cspsepsoc, void *svecsoc, double *residual_states, int *ksconv,
                   int *ksitmax) {
  (void)npxmax; (void)npymax; (void)npzmax; (void)key_natpri_out; (void)key_soc_calc;
  (void)key_pp_paw; (void)dijsoc; (void)ndisp; (void)ratio_diis; (void)latom; (void)natsoc;
  (void)ntyppp;
  (void)svalsoc; (void)cspsepsoc; (void)svecsoc; (void)xmax; (void)ymax; (void)zmax;
  API_BEGIN
  if (nso != 0 || ncol != 1)
    throw Error(5, "scf_diag: noncollinear spin and spin-orbit (nso=1 or ncol=2) are not part of this library");
  // scf_diag8f.F90:72-93
  bool l_rrb, l_rra, l_hsv, l_reo;
  const bool l_soc2nd = false;
  if (nrrz != 0) {
    const int az = std::abs(nrrz);
    l_rrb = (iscf % az == 0) && nrrz > 0;
    l_rra = (iscf % az == 0) && nrrz < 0;
    l_hsv = l_rra || l_soc2nd || (nrrz > 0 && (iscf + 1) % nrrz == 0);
  } else {
    l_rrb = l_rra = false;
    l_hsv = l_soc2nd;
  }
  l_reo = nchange > 0 ? ((iscf % nchange == 0) && !l_rra && !l_soc2nd) : false;
  if (nrc != 0) {
    BlochDiagArgs b{};
    b.nso = nso; b.ncol = ncol; b.iscf = iscf; b.ncgscf = ncgscf; b.nretcg = nretcg;
    b.natom = natom; b.num_spe = num_spe; b.neigmx = neigmx; b.nums = nums; b.nspv = nspv;
    b.numk = numk; b.nperi = nperi; b.northo = northo; b.nf = nf; b.num_list = num_list;
    b.num_ppcell = num_ppcell; b.nprjmx = nprjmx; b.ncpx = ncpx; b.ncpy = ncpy; b.ncpz = ncpz;
    b.key_cmpt = key_ortho_cmpt_innerproduct; b.key_nocmpt = key_ortho_nocmpt_innerproduct;
    b.key_in = key_natpri_in; b.key_inps = key_natpri_inps; b.nkscg = nkscg; b.nsdmax = nsdmax;
    b.nprecon_cg = nprecon_cg; b.l_rrb = l_rrb; b.l_rra = l_rra; b.l_reo = l_reo;
    b.ndiismax = ndiismax; b.nprecon_diis = nprecon_diis; b.l_hsv = l_hsv;
    b.eps_eig_diis = eps_eig_diis; b.alambda_diis = alambda_diis; b.alambda_min = alambda_min;
    b.alambda_max = alambda_max;
    b.dx = dx; b.dy = dy; b.dz = dz; b.epssd = epssd; b.skpxx = skpxx; b.skpyy = skpyy;
    b.skpzz = skpzz; b.dij = dij; b.veff = veff; b.vnlocp = vnlocp; b.sss = sss;
    b.indspe = indspe; b.natpri = natpri; b.naps = naps; b.nprj = nprj; b.natinf = natinf;
    b.lstvec2 = lstvec2; b.lstx = lstx; b.lsty = lsty; b.lstz = lstz; b.natx = natx;
    b.naty = naty; b.natz = natz; b.cspsep = cspsep; b.sveccm = sveccm; b.hsvcm = hsvcm;
    b.ssvcm = ssvcm; b.sval = sval; b.residual_states = residual_states; b.ksconv = ksconv;
    b.ksitmax = ksitmax;
    scf_diag_bloch(b);
    sync();
    return 0;
  }
  Grid g{ncpx, ncpy, ncpz, nf, nperi, dx, dy, dz};
  Ctx c(g);
  const long n = g.n(), n1 = (long)nprjmx * num_ppcell, nall = (long)neigmx * nums * numk;
  Arr vnl(vnlocp, (size_t)num_list * nprjmx * num_ppcell * D, Arr::In);
  Arr lst(lstvec2, (size_t)num_list * num_ppcell * I, Arr::In);
  Arr Dij(dij, (size_t)nprjmx * nprjmx * nums * ncol * natom * D, Arr::In);
  Arr Sss(sss, (size_t)nprjmx * nprjmx * num_spe * D, Arr::In);
  Arr Ve(veff, (size_t)n * nspv * D, Arr::In);
  Arr Rs(rspsep, (size_t)n1 * nall * D, Arr::InOut);
  Arr Ps(svecre, (size_t)n * nall * D, Arr::InOut);
  Arr Sp(ssvre, (size_t)n * nall * D, Arr::InOut);
  Arr Hp(hsvre, (size_t)n * neigmx * D, Arr::InOut);
  Arr Sv(sval, (size_t)nall * D, Arr::InOut);
  Arr Re(residual_states, (size_t)nall * D, Arr::InOut);
  setup(c, natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps,
        nprj, indspe, natinf, natpri, naps, vnl, lst, Dij.d(), Sss.d(), sss);
  *ksconv = 1;
  *ksitmax = 0;
  const bool block = param("kscg_mode", 0.0) >= 1.0;
  for (int nk = 1; nk <= numk; nk++) {
    for (int ns = 1; ns <= nums; ns++) {
      const size_t off = (size_t)n * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
      const size_t off1 = (size_t)n1 * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
      const size_t offs = (size_t)neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
      const int nsv = std::min(ns, nspv);
      const double *vs = Ve.d() + (size_t)n * (nsv - 1);
      Ctx &cc = c;
      const bool qz = cc.q_zero;
      if (l_rrb) {
        cc.q_zero = qz;
        Timed ph("phase_rr", 0.0);
        // block CG recomputes H psi from the rotated bands: hsvre is scratch here
        rayleigh_ritz(cc, neigmx, ns, vs, true, Ps.d() + off, Sp.d() + off, Rs.d() + off1, Hp.d(),
                      Sv.d() + offs, false, block);
      }
      if (iscf <= ncgscf || iscf % nretcg == 0) {
        KscgArgs a{neigmx, ns, northo, nkscg, nsdmax, nprecon_cg, key_ortho_cmpt_innerproduct,
                   key_ortho_nocmpt_innerproduct, epssd, vs, Ps.d() + off, Sp.d() + off, Hp.d(),
                   Rs.d() + off1, Sv.d() + offs, Re.d() + offs, ksconv, ksitmax};
        Timed ph("phase_cg", 0.0);
        // H psi is needed afterwards only by a Rayleigh-Ritz without its own H psi (l_rra); the
        // next scf_diag recomputes it (RR with l_rrb, or the block CG's first pass)
        if (block) kscg_block(cc, a, !l_rrb, l_rra);
        else kscg_sequential(cc, a);
      } else {
        DiisArgs a{neigmx, ns, ndiismax, nprecon_diis, northo, key_ortho_cmpt_innerproduct,
                   l_hsv ? 1 : 0, eps_eig_diis, alambda_diis, alambda_min, alambda_max, epssd,
                   Ve.d() + (size_t)n * (ns - 1), Ps.d() + off, Sp.d() + off, Hp.d(), Rs.d() + off1,
                   Sv.d() + offs, Re.d() + offs, ksconv, ksitmax};
        if (nspv == nums) ksdiis(cc, a);
      }
      if (l_rra)
        rayleigh_ritz(cc, neigmx, ns, vs, false, Ps.d() + off, Sp.d() + off, Rs.d() + off1, Hp.d(),
                      Sv.d() + offs, false);
    }
    if (l_reo)
      for (int ns = 1; ns <= nums; ns++) {
        Timed ph("phase_reorder", 0.0);
        const size_t off = (size_t)n * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
        const size_t off1 = (size_t)n1 * neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
        const size_t offs = (size_t)neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
        reorder(c, neigmx, Sv.d() + offs, Re.d() + offs, Ps.d() + off, Sp.d() + off, Rs.d() + off1,
                Hp.d(), false, block);  // (l_reo excludes l_rra: nothing reads hsvre after this)
      }
  }
  Rs.finish(); Ps.finish(); Sp.finish(); Hp.finish(); Sv.finish(); Re.finish();
  API_END
}

int kukan_scf_hartree(int nd, int ncpx_d, int ncpy_d, int ncpz_d, int nfh, int ndisp, int nperi,
                      int ncgres, int ncgmin, int ncgmax, double xmax, double ymax, double zmax,
                      double epsvh, const double *rho_aug_dense, const double *boundx,
                      const double *boundy, const double *boundz, double *vh_dense) {
  (void)ndisp;
  API_BEGIN
  if (nfh != 4 && nfh != 5 && nfh != 6)
    throw Error(6, "error in scf_hartree! This subroutine is for nfh ne 4, 5 or 6.");
  const long n = (long)ncpx_d * ncpy_d * ncpz_d;
  const size_t bx = nd ? 1 : (size_t)2 * nfh * ncpy_d * ncpz_d;
  const size_t by = nd ? 1 : (size_t)ncpx_d * 2 * nfh * ncpz_d;
  const size_t bz = nd ? 1 : (size_t)ncpx_d * ncpy_d * 2 * nfh;
  Arr rho(rho_aug_dense, n * D, Arr::In);
  Arr Bx(boundx, bx * D, Arr::In), By(boundy, by * D, Arr::In), Bz(boundz, bz * D, Arr::In);
  Arr vh(vh_dense, n * D, Arr::InOut);
  HartreeArgs a{nd, ncpx_d, ncpy_d, ncpz_d, nfh, nperi, ncgres, ncgmin, ncgmax, xmax, ymax, zmax,
                epsvh, rho.d(), Bx.d(), By.d(), Bz.d(), vh.d()};
  const int loops = hartree(a);
  lib().params["hartree_iterations"] = loops;
  vh.finish();
  API_END
}

int kukan_tools_correctpotential(int ncpx_d, int ncpy_d, int ncpz_d, double xmax, double ymax,
                                 double zmax, double ddx, double ddy, double ddz,
                                 double *vh_dense) {
  API_BEGIN
  const long n = (long)ncpx_d * ncpy_d * ncpz_d;
  Arr vh(vh_dense, n * D, Arr::InOut);
  correct_potential(n, ddx * ddy * ddz / (8.0 * xmax * ymax * zmax), vh.d());
  vh.finish();
  API_END
}

int kukan_scf_charge_smt_r(int neigmx, int nums, int nspv, int numk, int ncpx, int ncpy, int ncpz,
                           const double *fnele, const double *svecre, double *rhosmt) {
  API_BEGIN
  const long n = (long)ncpx * ncpy * ncpz, nall = (long)neigmx * nums * numk;
  Arr F(fnele, (size_t)nall * D, Arr::In);
  Arr Ps(svecre, (size_t)n * nall * D, Arr::In);
  Arr Rho(rhosmt, (size_t)n * nspv * D, Arr::Out);
  launch_fill(n * nspv, 0.0, Rho.d(), lib().stream);
  for (int ns = 1; ns <= nums; ns++) {
    const int nsr = std::min(ns, nspv);
    for (int nk = 1; nk <= numk; nk++) {
      const size_t b = (size_t)neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
      charge(n, neigmx, Ps.d() + n * b, F.d() + b, Rho.d() + (size_t)n * (nsr - 1));
    }
  }
  Rho.finish();
  API_END
}

int kukan_band_energy_r(int nperi, int neigmx, int nf, int ncpx, int ncpy, int ncpz, int num_list,
                        int natom, int num_spe, int num_ppcell, int nprjmx, int key_natpri_in,
                        int key_natpri_inps, const int *lstvec2, const int *nprj,
                        const int *natpri, const int *naps, const int *indspe, const int *natinf,
                        double xmax, double ymax, double zmax, const double *vnlocp,
                        const double *dij, const double *veff, const double *svecre,
                        double *eband) {
  API_BEGIN
  Grid g{ncpx, ncpy, ncpz, nf, nperi, 2.0 * xmax / ncpx, 2.0 * ymax / ncpy, 2.0 * zmax / ncpz};
  Ctx c(g);
  const long n = g.n();
  Arr vnl(vnlocp, (size_t)num_list * nprjmx * num_ppcell * D, Arr::In);
  Arr lst(lstvec2, (size_t)num_list * num_ppcell * I, Arr::In);
  Arr Dij(dij, (size_t)nprjmx * nprjmx * natom * D, Arr::In);
  Arr Ve(veff, (size_t)n * D, Arr::In);
  Arr Ps(svecre, (size_t)n * neigmx * D, Arr::In);
  Arr E(eband, (size_t)neigmx * D, Arr::Out);
  setup(c, natom, nprjmx, num_spe, num_ppcell, num_list, 1, key_natpri_in, key_natpri_inps, nprj,
        indspe, natinf, natpri, naps, vnl, lst, Dij.d(), nullptr, nullptr);
  band_energy(c, neigmx, Ve.d(), Ps.d(), E.d());
  E.finish();
  API_END
}

int kukan_energy_adjoint_r(int nperi, int neigmx, int nf, int ncpx, int ncpy, int ncpz,
                           int num_list, int natom, int num_spe, int num_ppcell, int nprjmx,
                           int key_natpri_in, int key_natpri_inps, const int *lstvec2,
                           const int *nprj, const int *natpri, const int *naps,
                           const int *indspe, const int *natinf, double xmax, double ymax,
                           double zmax, const double *vnlocp, const double *dvnlocp,
                           const double *dij, const double *svecre, const double *fnele,
                           double *dedv, double *dedr) {
  API_BEGIN
  Grid g{ncpx, ncpy, ncpz, nf, nperi, 2.0 * xmax / ncpx, 2.0 * ymax / ncpy, 2.0 * zmax / ncpz};
  Ctx c(g);
  const long n = g.n();
  const size_t tab = (size_t)num_list * nprjmx * num_ppcell;
  Arr vnl(vnlocp, tab * D, Arr::In);
  Arr dvnl(dvnlocp, 3 * tab * D, Arr::In);
  Arr lst(lstvec2, (size_t)num_list * num_ppcell * I, Arr::In);
  Arr Dij(dij, (size_t)nprjmx * nprjmx * natom * D, Arr::In);
  Arr Ps(svecre, (size_t)n * neigmx * D, Arr::In);
  Arr F(fnele, (size_t)neigmx * D, Arr::In);
  Arr Dv(dedv, (size_t)n * D, Arr::Out);
  Arr Dr(dedr, (size_t)3 * natom * D, Arr::Out);
  setup(c, natom, nprjmx, num_spe, num_ppcell, num_list, 1, key_natpri_in, key_natpri_inps, nprj,
        indspe, natinf, natpri, naps, vnl, lst, Dij.d(), nullptr, nullptr);
  std::vector<int> pri = host_ints(natpri, natom), isin(std::max(1, c.pl.nact));
  for (int a = 0; a < c.pl.nact; a++) isin[a] = pri[c.pl.act_h[5 * a + 3]] == key_natpri_in;
  int *isin_d = (int *)scratch("adj.isin", sizeof(int) * isin.size());
  HIPCK(hipMemcpyAsync(isin_d, isin.data(), sizeof(int) * isin.size(), hipMemcpyHostToDevice, c.s));
  energy_adjoint(c, neigmx, Ps.d(), F.d(), dvnl.d(), isin_d, Dv.d(), Dr.d());
  Dv.finish(); Dr.finish();
  API_END
}

int kukan_batch_create(void **handle) {
  API_BEGIN
  if (!handle) throw Error(1, "batch_create: null handle pointer");
  *handle = batch_create();
  API_END
}

int kukan_batch_add_system(void *handle, int n, double h, int nbands, int natom, int num_list,
                           const int *lstvec2, const int *natinf, const double *vnlocp,
                           const double *dij, const double *veff, const double *svecre,
                           int *index) {
  API_BEGIN
  const int i = batch_add(handle, n, h, nbands, natom, num_list, lstvec2, natinf, vnlocp, dij, veff,
                          svecre);
  if (index) *index = i;
  API_END
}

int kukan_batch_step(void *handle, int batched, int ncgmax, double epsvh) {
  API_BEGIN
  if (ncgmax < 1) throw Error(1, "batch_step: ncgmax must be positive");
  batch_step(handle, batched, ncgmax, epsvh);
  API_END
}

int kukan_batch_get(void *handle, int index, double *svecre, double *sval, double *vh,
                    int *hartree_iterations) {
  API_BEGIN
  batch_get(handle, index, svecre, sval, vh, hartree_iterations);
  API_END
}

int kukan_batch_destroy(void *handle) {
  API_BEGIN
  batch_destroy(handle);
  API_END
}

int kukan_force_eig_r(int nperi, int numk, int nums, int neigmx, int nf, int ncpx, int ncpy,
                      int ncpz, int num_list, int natom, int num_spe, int num_ppcell, int nprjmx,
                      int key_natpri_in, int key_natpri_inps, const int *lstvec2,
                      const int *latom, const int *lstx, const int *lsty, const int *lstz,
                      const int *nprj, const int *natpri, const int *naps, const int *indspe,
                      const int *natinf, double xmax, double ymax, double zmax,
                      const double *sss, const double *vnlocp, const double *dij,
                      const double *svecre, const double *sval, const double *fnele,
                      double *fatx, double *faty, double *fatz) {
  (void)latom; (void)lstx; (void)lsty; (void)lstz;
  API_BEGIN
  Grid g{ncpx, ncpy, ncpz, nf, nperi, 2.0 * xmax / ncpx, 2.0 * ymax / ncpy, 2.0 * zmax / ncpz};
  Ctx c(g);
  const long n = g.n(), nall = (long)neigmx * nums * numk;
  Arr vnl(vnlocp, (size_t)num_list * nprjmx * num_ppcell * D, Arr::In);
  Arr lst(lstvec2, (size_t)num_list * num_ppcell * I, Arr::In);
  Arr Dij(dij, (size_t)nprjmx * nprjmx * nums * natom * D, Arr::In);
  Arr Sss(sss, (size_t)nprjmx * nprjmx * num_spe * D, Arr::In);
  Arr Ps(svecre, (size_t)n * nall * D, Arr::In);
  Arr Sv(sval, (size_t)nall * D, Arr::In);
  Arr F(fnele, (size_t)nall * D, Arr::In);
  Arr Fx(fatx, (size_t)natom * D, Arr::Out), Fy(faty, (size_t)natom * D, Arr::Out),
      Fz(fatz, (size_t)natom * D, Arr::Out);
  setup(c, natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_natpri_in, key_natpri_inps,
        nprj, indspe, natinf, natpri, naps, vnl, lst, Dij.d(), Sss.d(), sss);
  std::vector<int> pri = host_ints(natpri, natom), isin(std::max(1, c.pl.nact));
  for (int a = 0; a < c.pl.nact; a++) isin[a] = pri[c.pl.act_h[5 * a + 3]] == key_natpri_in;
  int *isin_d = (int *)scratch("force.isin", sizeof(int) * isin.size());
  HIPCK(hipMemcpyAsync(isin_d, isin.data(), sizeof(int) * isin.size(), hipMemcpyHostToDevice, c.s));
  double *fa = c.buf("force.fa", (size_t)3 * natom);
  launch_fill(3L * natom, 0.0, fa, c.s);
  for (int nk = 1; nk <= numk; nk++)
    for (int ns = 1; ns <= nums; ns++) {
      const size_t b = (size_t)neigmx * ((ns - 1) + (size_t)nums * (nk - 1));
      force_eig(c, neigmx, ns, Ps.d() + n * b, Sv.d() + b, F.d() + b, isin_d, fa, nf);
    }
  HIPCK(hipMemcpyAsync(Fx.d(), fa, sizeof(double) * natom, hipMemcpyDeviceToDevice, c.s));
  HIPCK(hipMemcpyAsync(Fy.d(), fa + natom, sizeof(double) * natom, hipMemcpyDeviceToDevice, c.s));
  HIPCK(hipMemcpyAsync(Fz.d(), fa + 2L * natom, sizeof(double) * natom, hipMemcpyDeviceToDevice, c.s));
  Fx.finish(); Fy.finish(); Fz.finish();
  API_END
}

int kukan_totalenergy_eig(int nums, int ncol, int numk, int neigmx, const int *nwskk,
                          int nwskptot, double tf, const double *sval, const double *fnele,
                          double *eneel, double *eneth) {
  try {
    const int d2 = nums + 1 - ncol;
    std::vector<int> w = host_ints(nwskk, numk);
    std::vector<double> sv = host_doubles(sval, (size_t)neigmx * d2 * numk),
                        f = host_doubles(fnele, (size_t)neigmx * d2 * numk);
    double e = 0.0, t = 0.0;
    for (size_t k = 0; k < sv.size(); k++) e = e + f[k] * sv[k];
    for (int nk = 0; nk < numk; nk++) {
      const double tmpfac = (double)(nwskptot * nums) / (double)(2 * w[nk]);
      for (int ns = 0; ns < d2; ns++)
        for (int l = 0; l < neigmx; l++) {
          const double te = f[l + (size_t)neigmx * (ns + (size_t)d2 * nk)] * tmpfac;
          double t1 = 0.0, t2 = 0.0;
          if (te > 1.0e-8) t1 = std::log(te);
          if (1.0 - te > 1.0e-8) t2 = std::log(1.0 - te);
          t = t + (te * t1 + (1.0 - te) * t2) / tmpfac;
        }
    }
    *eneel = e;
    *eneth = t * tf;
    return 0;
  } catch (const std::exception &e) {
    lib().last_error = e.what();
    return 99;
  }
}

}  // extern "C"
