// kukan_internal.hpp — internal declarations of libkukan_hip.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace kukan {

// ----------------------------------------------------------------------------- errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
void hip_check(hipError_t e, const char *what, const char *file, int line);
#define HIPCK(x) ::kukan::hip_check((x), #x, __FILE__, __LINE__)
void blas_check(rocblas_status s, const char *what, const char *file, int line);
#define BLASCK(x) ::kukan::blas_check((x), #x, __FILE__, __LINE__)

// ----------------------------------------------------------------------------- library state
struct Lib {
  int device = -1;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  std::map<std::string, double> params;
  std::string last_error;
  // host pointer -> device mirror
  struct Mirror {
    void *dev;
    size_t bytes;
  };
  std::map<const void *, Mirror> bound;
  // grow-only scratch buffers keyed by tag
  std::map<std::string, std::pair<void *, size_t>> scratch;
  bool capturing = false;  // stream capture in progress: no timing events, no allocation
};
Lib &lib();
void ensure_init();
// once per device, with that device current: the device address of the fixed workspace of
// kukan_scf_hartree_noalloc (noalloc.hip)
void noalloc_setup(int device);
double param(const char *key, double dflt);
void *scratch(const std::string &tag, size_t bytes);  // device memory, reused across calls
void sync();

// Resolves an argument array to device memory for the duration of one entry-point call.
// Device pointers are used in place; bound host arrays use their mirror; other host
// arrays are copied in (if read) and back (if written) around the call.
class Arr {
 public:
  enum Intent { In = 1, Out = 2, InOut = 3 };
  Arr() = default;
  Arr(const void *p, size_t bytes, Intent it);
  Arr(const Arr &) = delete;
  Arr &operator=(const Arr &) = delete;
  Arr(Arr &&o) noexcept { *this = std::move(o); }
  Arr &operator=(Arr &&o) noexcept;
  ~Arr();
  void finish();  // copy back now (called on success paths)
  double *d() const { return (double *)dev_; }
  int *i() const { return (int *)dev_; }
  void *v() const { return dev_; }
  bool is_host_copy() const { return temp_; }

 private:
  const void *host_ = nullptr;
  void *dev_ = nullptr;
  size_t bytes_ = 0;
  Intent intent_ = In;
  bool temp_ = false;
};
// Copy a (host or device) int array to host.
std::vector<int> host_ints(const int *p, size_t n);
std::vector<double> host_doubles(const double *p, size_t n);
bool is_device_ptr(const void *p);

// ----------------------------------------------------------------------------- timing
// With param "timing" = 1, Timed regions record HIP events on the library stream; the
// accumulated device time per tag (and the algorithmic work the caller attributes to it)
// is read back through kukan_timing_get.  Regions whose tag starts with "phase_" (the SCF
// step's phases: Rayleigh-Ritz, CG, orthonormalisation, reorder, Hartree) are also roctx
// ranges, always, so that a profiler's marker trace cuts the timeline by phase.
class Timed {
 public:
  Timed(const char *tag, double work);
  ~Timed();

 private:
  const char *tag_;
  double work_;
  hipEvent_t a_ = nullptr, b_ = nullptr;
  bool range_ = false;
};
void timing_collect();
// adds work (no time, no launch) to a tag's accumulator while timing is on: a second measure
// of a timed launch, e.g. the MFMA flops a split-precision product issues beside its
// algorithmic flops
void timing_add_work(const char *tag, double work);
struct TimingAcc {
  double ms = 0.0, work = 0.0;
  long count = 0;
};
std::map<std::string, TimingAcc> &timing_table();

// ----------------------------------------------------------------------------- grid + nonlocal
struct Grid {
  int nx, ny, nz, nf, nperi;
  double dx, dy, dz;
  long n() const { return (long)nx * ny * nz; }
  double dv() const { return dx * dy * dz; }
};

// Active-atom tables and the inverse (grid point -> projector entries) map used for the
// deterministic owner-computes expansion.  Built on the device once per coarse call.
struct NonlocalPlan {
  int natom = 0, nprjmx = 0, num_list = 0, num_ppcell = 0, nums = 1, num_spe = 1;
  int nact = 0;                  // active atoms (natpri in {in, inps})
  const double *vnl = nullptr;   // vnlocp (num_list, nprjmx, num_ppcell), device
  const int *lst = nullptr;      // lstvec2 (num_list, num_ppcell), device, 1-based
  int *act = nullptr;            // per active atom: {iaps-1, natinf, iprj, na-1, ispe-1} x nact, device
  int *iprj_slot = nullptr;      // nprj for each ppcell slot (0 if unused), device
  int *row_ptr = nullptr;        // N+1
  int2 *ent = nullptr;           // {slot, k}
  long nent = 0;
  int *nzrows = nullptr;         // grid points with at least one entry, ascending
  long nnz = 0;
  std::vector<int> act_h;        // host copy of act
  int max_iprj = 0, max_natinf = 0;
};
void build_plan(NonlocalPlan &pl, const Grid &g, int natom, int nprjmx, int num_spe,
                int num_ppcell, int num_list, int nums, int key_in, int key_inps,
                const int *nprj, const int *indspe, const int *natinf, const int *natpri,
                const int *naps, const double *vnlocp_dev, const int *lstvec2_dev,
                const std::string &tag);

// ----------------------------------------------------------------------------- kernels (launchers)
// Hpsi for nb bands: out_b (=|+=) (-1/2 lap + veff) in_b + expansion with coef_b.
// in: interior arrays (wrap/zero halos per nperi) unless halo_in (explicit halo of width nf).
struct StencilCall {
  const Grid *g = nullptr;
  const double *in = nullptr;
  long ld_in = 0;      // band stride (elements)
  bool halo_in = false;
  double *out = nullptr;
  long ld_out = 0;
  const double *veff = nullptr;  // may be null
  int nb = 1;
  bool accumulate = false;
  const NonlocalPlan *pl = nullptr;  // expansion if non-null and coef non-null
  const double *coef = nullptr;      // (nprjmx, num_ppcell) per band, ld = ld_coef
  long ld_coef = 0;
  bool coef_band_fastest = false;    // coef laid out [slot][j][b] (see stencil_band_fastest_coef)
  double scale_lap = 1.0;            // multiplies the kinetic part (1 for H)
  double *dot_partial = nullptr;     // optional: per-block partial sums of in_b . out_b
  long dot_ld = 0;                   // band stride of dot_partial (0: number of parts)
  int nfh_coeffs = 0;                // if > 0: use this FD order (Hartree), no veff
  const int *skip = nullptr;         // device flag: the launch does nothing while *skip != 0
  // residual-preconditioner form (batched plans only): in = sign (in - eps_b in2)
  const double *in2 = nullptr, *eps = nullptr;
  double sign = 1.0;
  // combination form (tiled path only): the input band is ca_b in + cb_b in2, never stored;
  // out = ca_b (H in) + cb_b hin with hin = H in2 (by linearity), so the stencil loads only
  // in, and coef holds the projector coefficients of ca_b in alone; dot_partial (with the
  // combination at the centre) and dot_partial2 (same layout, in2 at the centre) receive the
  // partials of <ca in + cb in2|out> and <in2|out>.  With in2 = nullptr (symmetric form, for a
  // symmetric H; needs rp): dot_partial receives <ca in|out> (rp = <p|ca in>) and
  // dot_partial2 <ca in|hin>, and in2 is never read
  const double *ca = nullptr, *cb = nullptr;
  double *dot_partial2 = nullptr;
  const double *hin = nullptr;
  // timing tag of the stencil kernel (tiled path); work = input bands in + out, plus V_eff
  const char *tag = nullptr;
  // optional projections <p|in> (ld_rp per band) the coefficients were built from; with
  // them the projector part's partials come from projector space (one column at the
  // expansion's partial offset) instead of re-reading the bands.  rp2 = <p|in2> for
  // dot_partial2 of the combination form.
  const double *rp = nullptr, *rp2 = nullptr;
  long ld_rp = 0;
  // optional (tiled path): per-band max |in| and max |stencil part| as FP64 bits (atomic max;
  // zero them first), and per-band, per-work-group maxima of |projector part|
  // ([band][expand_nparts])
  unsigned long long *in_max = nullptr;
  double *norm_partial = nullptr;  // optional (tiled path): partials of <in|in>, as dot_partial
  unsigned long long *out_max = nullptr;
  double *expand_max = nullptr;
};

// Several systems' stencils as one persistent launch (device tables built once, so the
// launch can be captured in a graph).  res: 7-point preconditioner of the residual form.
struct StencilBatchPlan {
  void *sl = nullptr;
  long *start = nullptr;
  int nsys = 0, nf = 4, P = 2;
  long total = 0;
  bool res = false;
  size_t shm = 0;
};
StencilBatchPlan plan_stencil_batch(const std::vector<StencilCall> &calls, bool res, int P,
                                    const std::string &tag);
long stencil_batch_parts(const Grid &g, int nf, int nb, int P);
void launch_stencil_batch(const StencilBatchPlan &b, int grid, hipStream_t s);
void launch_stencil(const StencilCall &c, hipStream_t s);
// true when launch_stencil takes the tiled path, whose projector pass reads band-fastest coef
bool stencil_band_fastest_coef(const Grid &g, bool halo_in);
// partial-sum layout of launch_stencil's dot_partial: nb * nparts
int stencil_nparts(const Grid &g, int nf, int nb, bool comb = false);  // comb: combination form
// fused r = sign*(h - eps s), out = K r, optional per-band partials of <r|K r>
void launch_precond_resid_dot(const Grid &g, const double *h, const double *sp, long ld,
                              const double *eps, double sign, double *out, int nb,
                              double *dot_partial, hipStream_t s,
                              unsigned long long *out_max = nullptr);
// band update fused with the next residual preconditioner (tiled path, S = 1): per band
// xn = ca x + cb y, un = ha h + hb w into x_out, h_out (not x, h), out = K r with
// r = sign (un - eps xn); partials of <r|K r>, <xn|xn>, <xn|un> (update_precond_nparts(g, nb)
// per band at stride part_ld, which must be at least that), and max |K r| as FP64 bits
int update_precond_nparts(const Grid &g, int nb);
void launch_update_precond(const Grid &g, const double *x, const double *y, const double *h,
                           const double *w, long ld, const double *ca, const double *cb,
                           const double *ha, const double *hb, const double *eps, double sign,
                           double *x_out, double *h_out, double *out, int nb, double *part_r,
                           double *part_xx, double *part_xh, long part_ld,
                           unsigned long long *out_max, hipStream_t s);

void launch_halo_fill(double *v, int nperi, int na, int nb, int nc, int nf, int nf1, int nsnd,
                      hipStream_t s);
// out = K in (7-point smoother); halo_in: explicit halo width nf; else wrap/zero by nperi
void launch_precond(const Grid &g, const double *in, long ld_in, bool halo_in, int nf_halo,
                    double *out, long ld_out, int nb, hipStream_t s);
// out_b = K r_b with r_b = -(h_b - eps_b s_b) computed on the fly (fused residual).
void launch_precond_resid(const Grid &g, const double *h, const double *sp, long ld,
                          const double *eps, double sign, double *r_out, double *out, int nb,
                          hipStream_t s);

// <p|psi>: rp(i, slot, b) (=|+=) sum_k vnl(k,i,slot) psi_b(lst(k,slot)-1)
void launch_gather(const NonlocalPlan &pl, const double *psi, long ld, int nb, double *rp,
                   long ld_rp, bool accumulate, hipStream_t s);
// coef(j, slot, b) = sum_i rp(i, slot, b) X(i, j; atom) / dV, X = dij(:,:,ns,na) or sss(:,:,1,ispe)
void launch_coef(const NonlocalPlan &pl, const double *rp, long ld_rp, int nb, const double *X,
                 bool use_sss, int ns, double dvinv, double *coef, long ld_coef, hipStream_t s,
                 bool band_fastest = false);
// out_b(r) += sum_{entries of r} sum_j coef_b(j, slot) vnl(k, j, slot)
void launch_expand(const NonlocalPlan &pl, long n, const double *coef, long ld_coef, int nb,
                   double *out, long ld_out, hipStream_t s);
// Same over the compacted rows, 16 bands per thread; optional per-band partials of
// <psi_b|added part> written at part[b*part_ld + part_off + block].
int expand_nparts(const NonlocalPlan &pl);
// part[b*part_ld + part_off] = sum over active (slot, j) of coef(slot, j, b) * rp(slot, j, b):
// sum_r in_b(r) added_b(r) by linearity (rp = <p|in>, added = sum coef vnl), no grid pass.
// coef band-fastest [slot][j][b]; rp [b][slot][j] with ld_rp per band.
void launch_proj_dots(const NonlocalPlan &pl, const double *coef, int nb, const double *rp,
                      long ld_rp, double *part, long part_ld, int part_off, hipStream_t s);
void launch_expand_rows(const NonlocalPlan &pl, const double *coef, int nb, double *out,
                        long ld_out, const double *psi, long ld_psi, double *part, long part_ld,
                        int part_off, hipStream_t s,
                        const double *in2 = nullptr, const double *ca = nullptr,
                        const double *cb = nullptr, double *part2 = nullptr,
                        double *maxpart = nullptr);
// scale[b] = 2^(14 - e) with smax_b (FP64 bits) + max_i maxpart[b][i] < 2^e (maxpart may be null)
void launch_bound_scale(int nb, const unsigned long long *smax, const double *maxpart, int np,
                        double *scale, hipStream_t s);
// (coef band-fastest [slot][j][b]; with in2: the partials use ca_b psi + cb_b in2 as the band
// and part2 receives <in2|added>)

// batched reductions: res[b] = scale * sum_i x_b[i] * y_b[i] (deterministic two-stage);
// part: room for dot_nparts(n) * nb partial sums, or null for the library's own buffer
void launch_dots(long n, const double *x, long ldx, const double *y, long ldy, int nb,
                 double scale, double *res, hipStream_t s, double *part = nullptr);
int dot_nparts(long n);  // <= 1024
// res[b] = scale * sum over partials (nb x nparts) in fixed order
void launch_reduce_partials(const double *part, int nparts, int nb, double scale, double *res,
                            hipStream_t s);

// elementwise helpers (per-band scalars are device arrays)
void launch_axpby(long n, int nb, const double *a, double ca, const double *x, long ldx,
                  const double *b, double cb, const double *y, long ldy, double *out, long ldo,
                  hipStream_t s);  // out = (ca*a_b) x + (cb*b_b) y ; a/b may be null => 1
void launch_scale(long n, int nb, const double *f, double *x, long ldx, hipStream_t s);
// P = x + t_b P (t null: P = x); res1[b] = scale <sv_b|P_b>, res2[b] = scale <P_b|P_b>
void launch_pupdate_dots(long n, int nb, const double *x, const double *t, double *P, long ld,
                         const double *sv, double scale, double *res1, double *res2,
                         hipStream_t s);
// x = a_b x + c_b y, u = a2_b u + c2_b w (a2, c2 null: a, c); res1[b] = scale <x_b|x_b>,
// res2[b] = scale <x_b|u_b>.  x_out / u_out (null: x / u) receive the new x and u.
void launch_pair_update_dots(long n, int nb, const double *a, const double *c, double *x,
                             const double *y, double *u, const double *w, long ld, double scale,
                             double *res1, double *res2, hipStream_t s,
                             const double *a2 = nullptr, const double *c2 = nullptr,
                             double *x_out = nullptr, double *u_out = nullptr);
void launch_copy2d(long n, int nb, const double *x, long ldx, double *y, long ldy, hipStream_t s);
void launch_fill(long n, double v, double *x, hipStream_t s);
void launch_fill_normal(double *x, long n, unsigned long long seed, hipStream_t s);
// y(:, dst[j]) = x(:, src[j]) for j < cnt (device index lists; null = identity)
void launch_copy_cols(long n, int cnt, const int *src, const int *dst, const double *x, long ldx,
                      double *y, long ldy, hipStream_t s);
void launch_permute_cols(long n, int nb, const int *perm_dev, const double *x, long ldx,
                         double *y, long ldy, hipStream_t s);  // y(:,perm[j]) = x(:,j)
// sign-canonicalise eigenvector columns (largest |.| positive, lowest index on ties)
void launch_canon_sign(int n, double *z, hipStream_t s);

}  // namespace kukan
