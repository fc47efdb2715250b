// solvers.hpp — device-pointer implementations of the coarse routines.
#pragma once
#include "gemm_mixed.hpp"
#include "kukan_internal.hpp"

namespace kukan {

void fd_coefficients(int nf, double *acfd);

// Per-call context: grid, nonlocal plan and the device tables of one coarse call.
struct Ctx {
  explicit Ctx(const Grid &g);
  Grid g;
  NonlocalPlan pl;
  hipStream_t s;
  const double *dij = nullptr, *sss = nullptr;
  bool q_zero = true;  // S = 1 (all sss entries zero): S psi == psi
  int dij_sym = -1;    // dij exactly symmetric: 1 / 0, -1 = not yet checked
  bool dij_symmetric();
  double *buf(const char *tag, size_t count);
  int *flag(int i = 0);
  void clear_flag(int i = 0);
  int read_flag(int i = 0);
  void proj(const double *psi, long ld, int nb, double *rp, long ld_rp);
  // rp_exact: rp was just computed from psi, so eps_out's projector part may be taken from
  // projector space (sum coef rp) instead of a second pass over the bands.  pscale_out: split
  // scales 2^(14 - e) of psi from its column maxima; hscale_out: the same for out from a bound
  // on its column maxima (both from the same pass on the tiled path, else from passes of
  // their own)
  // norm_out: dV <psi|psi> per band (from the same pass on the tiled path)
  void hpsi(int ns, const double *veffs, const double *psi, long ld, int nb, const double *rp,
            long ld_rp, double *out, long ld_out, bool accumulate, double *eps_out,
            bool rp_exact = false, double *pscale_out = nullptr, double *hscale_out = nullptr,
            double *norm_out = nullptr);
  // H p for p = ca_b x + cb_b y, p never stored (tiled path): out = H p = ca_b H x + hy with
  // hy = H y, pap = dV <p|H p>, sap = dV <y|H p>.  rpx = ca_b <proj|x> (the projector
  // coefficients of ca_b x), rp = <proj|p>; with rpy = <proj|y> the projector parts of pap and
  // sap come from projector space.  With eh = dV <y|H y> as well and dij symmetric, sap is
  // formed as dV <H y|ca x> + cb eh (H symmetric) and y is not read.
  void hpsi_comb(int ns, const double *veffs, const double *x, const double *y, const double *hy,
                 long ld, int nb, const double *ca, const double *cb, const double *rpx,
                 const double *rp, long ld_rp, double *out, double *pap, double *sap,
                 const double *rpy = nullptr, const double *eh = nullptr);
  void spsi(const double *psi, long ld, int nb, const double *rp, long ld_rp, double *out,
            long ld_out);
  void dots(const double *x, long ldx, const double *y, long ldy, int nb, double *res);
};

struct KscgArgs {
  int neigmx, ns, northo, nkscg, nsdmax, nprecon_cg, key_cmpt, key_nocmpt;
  double epssd;
  const double *veffs;
  double *psi, *spsi, *hpsi, *rsp, *sval, *resid;
  int *ksconv, *ksitmax;
};
struct DiisArgs {
  int neigmx, ns, ndiismax, nprecon, northo, key_cmpt, l_hsv;
  double eps_eig_diis, alambda_diis, alambda_min, alambda_max, epssd;
  const double *veffs;
  double *psi, *spsi, *hpsi, *rsp, *sval, *resid;
  int *ksconv, *ksitmax;
};
struct HartreeArgs {
  int nd, nx, ny, nz, nfh, nperi, ncgres, ncgmin, ncgmax;
  double xmax, ymax, zmax, epsvh;
  const double *rho, *boundx, *boundy, *boundz;
  double *vh;
};

// C(M1 x M2) = alpha A^T B + beta C (A: K x M1, B: K x M2) and Y(R x M2) = alpha X C + beta Y
// (X: R x K, Y must not alias X), all column-major on the library stream; FP64 or split-FP16
// per params "precision" / "mixed_min_rows" / "mixed_terms".
// Column scales sa/sb/sx may be passed in (from mixed_scale) when an operand is reused;
// null means "compute here".  mixed_scale returns null in FP64 mode.
// tri (split precision): 1 / 2 computes only the tiles on and below / above the diagonal
void subspace_tn(int M1, int M2, long K, double alpha, const double *A, long lda,
                 const double *B, long ldb, double beta, double *C, long ldc,
                 const double *sa = nullptr, const double *sb = nullptr,
                 const SplitBands *pa = nullptr, int tri = 0);
// colsq (split precision, beta = 0): dv sum_i Y(i,j)^2 per column of the result
void subspace_nn(long R, int K, int M2, double alpha, const double *X, long ldx, const double *C,
                 long ldc, double beta, double *Y, long ldy, const double *sx = nullptr,
                 const SplitBands *px = nullptr, double *colsq = nullptr, double dv = 1.0,
                 bool c_upper = false);
const double *mixed_scale(const char *tag, const double *X, long ldx, long rows, int cols);

// fp64: FP64 pass also when the subspace products are split-precision
void cholqr(Ctx &c, int M, double *psi, double *spsi, double *rsp, double *hpsi,
            const double *norms = nullptr, bool fp64 = false);
void gs_one(Ctx &c, int l, double *psi, double *spsi, double *rsp);
void ortho_r_01(Ctx &c, int M, int l_inp, int nflag, int key_cmpt, double *psi, double *spsi,
                double *rsp);
// h_scratch: hpsi's content after the call is not used by the caller (the block CG
// recomputes H psi), so the split-precision path rotates psi in place without filling it
void rayleigh_ritz(Ctx &c, int M, int ns, const double *veffs, bool l_rrb, double *psi,
                   double *spsi, double *rsp, double *hpsi, double *sval, bool keep_h,
                   bool h_scratch = false);
void kscg_sequential(Ctx &c, const KscgArgs &a);
// keep_h: leave H psi of the returned bands in a.hpsi (kscg_r's hsvre); false when the caller
// recomputes it (scf_diag's block path), which saves rotating it in the closing Cholesky-QR
void kscg_block(Ctx &c, const KscgArgs &a, bool ortho_first, bool keep_h = true);
void ksdiis(Ctx &c, const DiisArgs &a);
// reorderstates8f.F90:8-124.  Default: hpsi is the work array and ends as the permuted S psi
// (the reference's hsvre).  move_h (block path): hpsi holds H psi and is permuted with psi;
// h_scratch (block path of scf_diag): hpsi's content is not used afterwards.  In both block
// forms only displaced bands move.
void reorder(Ctx &c, int M, double *sval, double *resid, double *psi, double *spsi, double *rsp,
             double *hpsi, bool move_h = false, bool h_scratch = false);
// Buffers of one Hartree solve carved from a caller's workspace instead of the library's
// own allocations: 5 grid vectors and the small arrays (see hartree_work_bytes).
struct HartreeWork {
  char *base;
  size_t bytes;
};
// bytes a Hartree solve on a grid of n points needs in a HartreeWork (0 if the grid's
// partial-sum counts are unknown before set-up: never for the tiled stencil)
size_t hartree_work_bytes(const Grid &g, int nfh);
// ws: run on the caller's workspace with eager launches (no graph, no allocation)
int hartree(const HartreeArgs &a, const HartreeWork *ws = nullptr);
void correct_potential(long n, double dvol_over_omega, double *vh);
void charge(long n, int M, const double *psi, const double *f, double *rho);
// e_l = dV <psi_l|H|psi_l> (H = T + veff + nonlocal with dij)
void band_energy(Ctx &c, int M, const double *veff, const double *psi, double *eband);
// adjoint of E = sum_l f_l e_l: dedv = dE/dveff (n), dedr = nonlocal dE/dR (3 x natom) from the
// projector derivative tables dvnl (3 tables of vnlocp's shape, d vnlocp / d R_x,y,z)
// batching pattern (batch.hip)
void *batch_create();
void batch_destroy(void *h);
int batch_add(void *h, int n, double hh, int nbands, int natom, int num_list, const int *lstvec2,
              const int *natinf, const double *vnlocp, const double *dij, const double *veff,
              const double *psi);
void batch_step(void *h, int batched, int ncgmax, double epsvh);
void batch_get(void *h, int i, double *psi, double *sval, double *vh, int *iters);
void energy_adjoint(Ctx &c, int M, const double *psi, const double *f, const double *dvnl,
                    const int *is_in_dev, double *dedv, double *dedr);
void force_eig(Ctx &c, int M, int ns, const double *psi, const double *sval, const double *fnele,
               const int *is_in_dev, double *fa, int nf);

}  // namespace kukan
