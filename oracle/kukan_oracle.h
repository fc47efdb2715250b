/*
 * kukan_oracle.h — FP64 CPU restatement of the RSPACE kukan v8f Kohn-Sham hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for libkukan_hip.so: only the
 * tests, the smoke check and the CPU-baseline leg of bench.py load it.  It is never part
 * of the product path.
 *
 * Each orc_<routine> has exactly the argument list of kukan_<routine> in
 * include/kukan_hip.h, which in turn is the reference Fortran routine's list.  Arrays are
 * Fortran-ordered (column-major, first index fastest); lstvec2 stays 1-based.
 * Routines return 0 on success and a nonzero code (message via orc_last_error()) where the
 * reference calls stopp().
 */
#ifndef KUKAN_ORACLE_H
#define KUKAN_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

const char *orc_last_error(void);
int orc_num_threads(void);

int orc_tools_definefdcoef(int nf, double *acfd);
int orc_overlap_finitedifference_r(int nperi, int ncpa, int ncpb, int ncpc, int nf, int nf1,
                                   int nsnd, double *v_cc);
int orc_kslaplacian_r(int ncpx, int ncpy, int ncpz, int neigmx, int nf, double dx, double dy,
                      double dz, const double *vre, const double *veff, int l, double *avre);
int orc_ksprecondition(int ncpx, int ncpy, int ncpz, int nf, int ncol, const double *vre,
                       double *avre);
int orc_nonlocaloperation_r_01(int natom, int nprjmx, int num_spe, int num_ppcell, int num_list,
                               int ncpx, int ncpy, int ncpz, int key_natpri_in,
                               int key_natpri_inps, const int *nprj, const int *indspe,
                               const int *natinf, const int *natpri, const int *naps,
                               const int *lstvec2, const double *vnlocp, const double *vre2,
                               double *rpsep);
int orc_nonlocaloperation_r_02(int nhs, int natom, int num_spe, int nums, int nprjmx, int ns,
                               int num_list, int num_ppcell, int ncpx, int ncpy, int ncpz,
                               int key_natpri_in, int key_natpri_inps, double dx, double dy,
                               double dz, const int *nprj, const int *indspe, const int *natinf,
                               const int *lstvec2, const int *natpri, const int *naps,
                               const double *vnlocp, const double *rpsep, const double *dijsss,
                               double *avre, double *avr);
int orc_kscg_r(int natom, int num_spe, int neigmx, int nums, int nspv, int numk, int nperi,
               int northo, int ns, int nk, int nf, int nkscg, int nsdmax, int nprecon_cg,
               int nprjmx, int num_list, int num_ppcell, int ndisp, int ncpx, int ncpy, int ncpz,
               int key_ortho_nocmpt_innerproduct, int key_ortho_cmpt_innerproduct,
               int key_natpri_in, int key_natpri_inps, int key_natpri_out, double dx, double dy,
               double dz, double epssd, const int *indspe, const int *natpri, const int *naps,
               const int *nprj, const int *natinf, const int *lstvec2, const int *latom,
               const int *ntyppp, const double *dij, const double *veff, const double *vnlocp,
               const double *sss, int *ksconv, int *ksitmax, double *sval, double *rspsep,
               double *svecre, double *ssvre, double *residual_states, double *hsvre);
int orc_ksdiis_r(int l_hsv, int natom, int num_spe, int neigmx, int nums, int nspv, int numk,
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
                 double *residual_states, double *hsvre);
int orc_orthogonalization_r(int natom, int neigmx, int nprjmx, int nums, int numk, int num_ppcell,
                            int num_spe, int northo, int num_list, int nflag, int ndisp, int ncpx,
                            int ncpy, int ncpz, int key_ortho_cmpt_innerproduct,
                            int key_natpri_in, int key_natpri_inps, double dx, double dy,
                            double dz, const int *nprj, const int *indspe, const int *natpri,
                            const int *naps, const int *natinf, const int *lstvec2,
                            const int *latom, const int *ntyppp, const double *vnlocp,
                            const double *sss, double *svecre, double *ssvre, double *rspsep);
int orc_orthogonalization_r_01(int natom, int neigmx, int nprjmx, int nums, int numk, int nk,
                               int ns, int l_inp, int num_list, int num_ppcell, int num_spe,
                               int nflag, int ndisp, int ncpx, int ncpy, int ncpz,
                               int key_ortho_cmpt_innerproduct, int key_natpri_in,
                               int key_natpri_inps, double dx, double dy, double dz,
                               const int *nprj, const int *indspe, const int *natpri,
                               const int *naps, const int *natinf, const int *lstvec2,
                               const int *latom, const double *vnlocp, const double *sss,
                               double *svecre, double *ssvre, double *rspsep);
int orc_rayleighritz_r(int natom, int num_ppcell, int num_spe, int nprjmx, int neigmx, int numk,
                       int nums, int nspv, int num_list, int ncpx, int ncpy, int ncpz, int nperi,
                       int nk, int ns, int nf, int ndisp, int key_natpri_in, int key_natpri_inps,
                       double dx, double dy, double dz, const int *nprj, const int *natinf,
                       const int *lstvec2, const int *indspe, const int *natpri, const int *naps,
                       const int *ntyppp, const double *dij, const double *veff,
                       const double *vnlocp, int l_rrb, double *rspsep, double *svecre,
                       double *ssvre, double *hsvre, double *sval);
int orc_reorderstates(int nrc, int num_ppcell, int nprjmx, int neigmx, int nums, int ncol,
                      int numk, int ncpx, int ncpy, int ncpz, int ns1, int nk, double *sval,
                      double *residual_states, double *rspsep, void *cspsep, double *svecre,
                      void *sveccm, double *ssvre, void *ssvcm, double *hsvre, void *hsvcm);
int orc_scf_diag(int nrc, int nso, int nprjmx, int num_ppcell, int num_spe, int iscf,
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
                 int *ksitmax);
int orc_scf_hartree(int nd, int ncpx_d, int ncpy_d, int ncpz_d, int nfh, int ndisp, int nperi,
                    int ncgres, int ncgmin, int ncgmax, double xmax, double ymax, double zmax,
                    double epsvh, const double *rho_aug_dense, const double *boundx,
                    const double *boundy, const double *boundz, double *vh_dense);
int orc_tools_correctpotential(int ncpx_d, int ncpy_d, int ncpz_d, double xmax, double ymax,
                               double zmax, double ddx, double ddy, double ddz, double *vh_dense);
int orc_scf_charge_smt_r(int neigmx, int nums, int nspv, int numk, int ncpx, int ncpy, int ncpz,
                         const double *fnele, const double *svecre, double *rhosmt);
int orc_force_eig_r(int nperi, int numk, int nums, int neigmx, int nf, int ncpx, int ncpy,
                    int ncpz, int num_list, int natom, int num_spe, int num_ppcell, int nprjmx,
                    int key_natpri_in, int key_natpri_inps, const int *lstvec2, const int *latom,
                    const int *lstx, const int *lsty, const int *lstz, const int *nprj,
                    const int *natpri, const int *naps, const int *indspe, const int *natinf,
                    double xmax, double ymax, double zmax, const double *sss,
                    const double *vnlocp, const double *dij, const double *svecre,
                    const double *sval, const double *fnele, double *fatx, double *faty,
                    double *fatz);
int orc_band_energy_r(int nperi, int neigmx, int nf, int ncpx, int ncpy, int ncpz, int num_list,
                      int natom, int num_spe, int num_ppcell, int nprjmx, int key_natpri_in,
                      int key_natpri_inps, const int *lstvec2, const int *nprj, const int *natpri,
                      const int *naps, const int *indspe, const int *natinf, double xmax,
                      double ymax, double zmax, const double *vnlocp, const double *dij,
                      const double *veff, const double *svecre, double *eband);
int orc_energy_adjoint_r(int nperi, int neigmx, int nf, int ncpx, int ncpy, int ncpz,
                         int num_list, int natom, int num_spe, int num_ppcell, int nprjmx,
                         int key_natpri_in, int key_natpri_inps, const int *lstvec2,
                         const int *nprj, const int *natpri, const int *naps, const int *indspe,
                         const int *natinf, double xmax, double ymax, double zmax,
                         const double *vnlocp, const double *dvnlocp, const double *dij,
                         const double *svecre, const double *fnele, double *dedv, double *dedr);
int orc_band_sd_step(int n, double h, int nbands, int natom, int num_list, const int *lstvec2,
                     const int *natinf, const double *vnlocp, const double *dij,
                     const double *veff, double *svecre, double *sval);
int orc_totalenergy_eig(int nums, int ncol, int numk, int neigmx, const int *nwskk, int nwskptot,
                        double tf, const double *sval, const double *fnele, double *eneel,
                        double *eneth);

/* Symmetric eigensolver used wherever the reference calls dsyev: Householder
 * tridiagonalisation + implicit QL (tred2/tql2 of Wilkinson & Reinsch, Handbook for
 * Automatic Computation II, 1971).  a(n,n) column-major, lower triangle read; on exit the
 * columns of a are eigenvectors for ascending w, each signed so that its component of
 * largest magnitude (lowest index on ties) is positive. */
int orc_dsyev(int n, double *a, double *w);

/* complex (Bloch) path: kukan_oracle_bloch.c.  Complex arrays are complex*16, interleaved
 * (re, im) doubles. */
int orc_overlap_finitedifference_c(int nperi, int ncpa, int ncpb, int ncpc, int nf, int nf1,
                                   int nsnd, int ncol, double *v_cc);
int orc_kslaplacian_c(int ncpx, int ncpy, int ncpz, int neigmx, int ncol, int nvef, int nf,
                      double dx, double dy, double dz, double xmax, double ymax, double zmax,
                      double skpx, double skpy, double skpz, double *vcm, const double *veff,
                      int l, double *avcm, double *workc);
int orc_ksprecondition_c(int ncpx, int ncpy, int ncpz, int nf, int ncol, const double *vcm,
                         double *avcm);
int orc_nonlocaloperation_c_01(int natom, int nprjmx, int num_spe, int num_ppcell, int num_list,
                               int neigmx, int ncol, int l, int ncpx, int ncpy, int ncpz,
                               int key_natpri_in, int key_natpri_inps, double dx, double dy,
                               double dz, double skpx, double skpy, double skpz, const int *nprj,
                               const int *indspe, const int *natinf, const int *natpri,
                               const int *naps, const int *lstvec2, const int *lstx,
                               const int *lsty, const int *lstz, const int *natx, const int *naty,
                               const int *natz, const double *vnlocp, const double *vcm2,
                               double *cpsep, double *vcccm);
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
                               const double *dijsoc, double *avcm, double *vcccm, double *avc);
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
                               const double *sss, double *sveccm, double *ssvcm, double *cspsep);
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
                            double *ssvcm, double *cspsep);
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
               double *sveccm, double *ssvcm, double *residual_states, double *hsvcm);
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
                       double *ssvcm, double *hsvcm, double *sval);
int orc_reorderstates_c(int num_ppcell, int nprjmx, int neigmx, int nums, int ncol, int ncpx,
                        int ncpy, int ncpz, int ns1, int nk, double *sval, double *residual_states,
                        double *cspsep, double *sveccm, double *ssvcm, double *hsvcm);
int orc_scf_charge_smt_c(int neigmx, int nums, int ncol, int nspv, int numk, int ncpx, int ncpy,
                         int ncpz, const double *fnele, const double *sveccm, double *rhosmt);
/* Hermitian eigensolver (cyclic Jacobi; upper triangle read), test helper */
int orc_zheev(int n, double *a, double *w);
/* kukan_oracle_ckpt.c: checkpoint files */
int orc_ckpt_path(const char *chdir, const char *chnam, int nr1, int nr2, char *out, int len);
int orc_write_wf(int nr1In, int nr2In, const char *chdir, int nrc, int nums, int ncol, int neigmx,
                 int numk, int ncpx, int ncpy, int ncpz, int nrxIn, int nryIn, int nrzIn, int ksym,
                 const double *svecre, const double *sveccm, const double *sval);
int orc_read_wf(int nr1In, int nr2In, const char *chdir, int nrc, int nums, int ncol, int neigmx,
                int numk, int ncpx, int ncpy, int ncpz, int ksym, int key_ksym_inv2, double *svecre,
                double *sveccm, double *sval);
int orc_read_rho(int nrIn, const char *chdir, int nevhist, int nspv, int ncpx, int ncpy, int ncpz,
                 int nradmx, int npoint, int num_atcell, int *lbrydn, int *na_atcell,
                 int *natcell_inf, double *rhosmt, double *rhotrur, double *rhosmtr);
int orc_read_vht(int nrIn, const char *chdir, int nevhist, int ncpx, int ncpy, int ncpz, int nmesh,
                 double *vh_dense);
int orc_write_rhovht(int nrIn, const char *chdir, int nrho, int nele, int npre, int nevhist, int nspv,
                     int ncpx, int ncpy, int ncpz, int nmesh, int nradmx, int npoint, int num_atcell,
                     int nrxIn, int nryIn, int nrzIn, int natom, int key_natpri_in, const int *natpri,
                     const double *rhosmt_i, const double *rhotrur_i, const double *rhosmtr_i,
                     const double *rhosmt_o, const double *rhotrur_o, const double *rhosmtr_o,
                     const double *vh_dense);
int orc_filenumbers_rhovht(const char *chdir, int *nrrho, int *nrpot);
/* kukan_oracle_trans.c: potential-assembly grid transfers */
int orc_overlap_interpolation(int ndisp, int nperi, int ncpa, int ncpb, int ncpc, int nf, int nf1,
                              double *v_cc, int nflag);
int orc_interpolation(int ncpx, int ncpy, int ncpz, int nmesh, double *v_dd, double *v_cc, int nperi,
                      int nf, int is, int ndisp);
int orc_trans_c2d_smtcharge(int nmesh, int nspv, int nperi, int ndisp, int nf, int ncpx, int ncpy, int ncpz,
                            const double *rhosmt, double *rho_dense);
int orc_trans_d2c_smtcharge(int nmesh, int nperi, int ndisp, int nf, int ncpx, int ncpy, int ncpz,
                            const double *rho_aug_dense, double *rho_coarse);
int orc_trans_d2c_vh(int natom, int num_spe, int nmesh, int nperi, int ndisp, int nf, int ncpx,
                     int ncpy, int ncpz, double xmax, double ymax, double zmax, const int *indspe,
                     const double *atmpole, const double *cp, const double *atx, const double *aty,
                     const double *atz, double *vh_coarse, const double *vh_dense);
int orc_trans_d2c_vxc(int nmesh, int nspv, int nperi, int ndisp, int nf, int ncpx, int ncpy, int ncpz,
                      double *vx, const double *vx_dense);
int orc_scf_rhoaugdense(int key_natpri_inps, int key_pp_paw, int natom, int num_spe, int nspv,
                        int ncpx_d, int ncpy_d, int ncpz_d, int num_list_d, int num_ppcell_d,
                        const int *indspe, const int *natprid, const int *napsd, const int *natinfd,
                        const int *ntyppp, const int *lstvecd2, const double *rho_dense,
                        const double *rhoaug3d, double *rho_aug_dense);
int orc_scf_veff(int nspv, int ncpx, int ncpy, int ncpz, double xmax, double ymax, double zmax,
                 double dx, double dy, double dz, double biasx, double biasy, double biasz,
                 const double *vloc_coarse, const double *vh_coarse, const double *vx, double *veff);
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
                 double *hsvcm);
/* kukan_oracle_pp.c: projector tables, list vectors and forces */
int orc_pseudocalc_01a(int nfdg, int npmesh, double *wxyz);
int orc_pseudocalc_01b(int nperi, int natom, int num_spe, int nradmx, int nmesh, int nfdg,
                       int ndisp, int num_atcell, int num_ppcell, int num_ppcell_d, int num_list,
                       int num_list_d, int nxmax, int nymax, int nzmax, int ncpx, int ncpy,
                       int ncpz, int npxmax, int npymax, int npzmax, int key_natpri_in,
                       int key_natpri_inps, int key_natpri_out, double psctoff, double xmax,
                       double ymax, double zmax, double dx, double dy, double dz,
                       const int *indspe, const int *nradct, const double *radial,
                       const double *atx, const double *aty, const double *atz, int *natx,
                       int *naty, int *natz, int *ndatx, int *ndaty, int *ndatz, int *lstx,
                       int *lsty, int *lstz, int *lstdx, int *lstdy, int *lstdz, int *natinf,
                       int *natinfd, int *natinfd_vloc, int *lstvec2, int *lstvecd2, int *natpri,
                       int *natprid, int *natpri_inf, int *naps, int *napsd);
int orc_pseudocalc_07(int natom, int num_spe, int npmesh, int nprjmx, int nfdg, int nfiltyp,
                      int num_list, int num_ppcell, int nradmx, int nqmx, int nxmax, int nymax,
                      int nzmax, int npxmax, int npymax, int npzmax, int nxspd, int nyspd,
                      int nzspd, int key_natpri_in, int key_natpri_inps, double eps,
                      double psftrad, double psctoff, double filpp, double dx, double dy,
                      double dz, const int *indspe, const int *nprj, const int *nradct,
                      const int *natpri, const int *naps, const int *natinf, const int *nqct,
                      const int *natx, const int *naty, const int *natz, const int *lstx,
                      const int *lsty, const int *lstz, const double *wxyz, const double *radial,
                      const double *coef, const double *atx, const double *aty,
                      const double *atz, double *vnlocp, double *vtmp_pp, double *vtmp_pp1,
                      double *vtmp_pp2, double *vnlspd);
int orc_force(int natom, int num_spe, int num_atcell, int num_ppcell_d, int num_list_d, int nqmx,
              int nradmx, int npoint, int nmesh, int nspv, int nperi, int nint1dmax, int nzmax,
              int jelcalc, int ncpx, int ncpy, int ncpz, int ncpx_d, int ncpy_d, int ncpz_d,
              int new_pwx, int new_pwy, int new_pwz, int new_rsx, int new_rsy, int new_rsz,
              int key_jel_calc, int key_natpri_in, int key_natpri_inps, const int *indspe,
              const int *natpri, const int *natpri_inf, const int *nradct, const int *natprid,
              const int *napsd, const int *natinfd, const int *natinfd_vloc,
              const int *lstvecd2, const int *nqctpcc, const int *lstdx, const int *lstdy,
              const int *lstdz, double veta, double psctoff, double psftrad,
              const double *radial, const double *dradial, const double *point,
              const double *wt, double xmax, double ymax, double zmax, double biasx,
              double biasy, double biasz, const double *cp, const double *coef,
              const double *atx, const double *aty, const double *atz,
              const double *dvlocdx_scw, const double *dvlocdy_scw, const double *dvlocdz_scw,
              const double *dvlocdx_hdp, const double *dvlocdy_hdp, const double *dvlocdz_hdp,
              const double *rhotrur, const double *rhosmtr, const double *rhoaugr,
              const double *drhoaug3ddx, const double *drhoaug3ddy, const double *drhoaug3ddz,
              const double *rho_coarse, const double *rho_aug_dense, const double *vloc_dense,
              const double *vh_dense, const double *vx_dense, double chrjel, double strjel,
              double endjel, const int *ndatx, const int *ndaty, const int *ndatz,
              double *fatx, double *faty, double *fatz, double *fjel);

#ifdef __cplusplus
}
#endif
#endif
