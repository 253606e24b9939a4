/*
 * er_oracle.h -- FP64 CPU restatement of ERmod erdst's per-frame energy and
 * distribution path.  TEST INFRASTRUCTURE ONLY: the oracle is the checker for
 * the HIP path and the timed host-core baseline.  It is never linked into,
 * called by, or shipped with the product library.
 *
 * Conventions (chosen to match the Fortran memory layout of the reference):
 *   positions   double pos[n][3]                (= sitepos(3,n) column-major)
 *   cell        double cell[9], (i,j) at i+3*j  (= cell(3,3) column-major;
 *                                                column j is box vector j)
 *   grids       index x1 + K1*(x2 + K2*x3)      (= cnvslt(0:K1-1,0:K2-1,0:K3-1))
 *   uvengy      double [nins][nmol]             (= uvengy(slvmax,maxdst))
 *   ecorr       double [ermax][ermax]           (symmetric)
 * All indices are 0-based; the reference's 1-based quantities are shifted.
 */
#ifndef ER_ORACLE_H
#define ER_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ljswitch values, erdst/engmain.F90:358-359 */
enum { ORC_LJSWT_POT_CHM = 0, ORC_LJSWT_POT_GMX = 1, ORC_LJSWT_FRC_CHM = 2, ORC_LJSWT_FRC_GMX = 3 };
/* slttype values, erdst/engmain.F90:362 */
enum { ORC_SLT_SOLN = 1, ORC_SLT_REFS_RIGID = 2, ORC_SLT_REFS_FLEX = 3 };

typedef struct {
    int natom;                 /* atoms of the physical system                   */
    int nmol;                  /* molecules of the physical system (= slvmax)    */
    const int *mol_begin;      /* [nmol+1], 0-based first atom of each molecule  */
    const double *charge;      /* [natom], already x 18.22261721                 */
    const int *ljtype;         /* [natom], 0-based                               */
    int ntype;                 /* ljtype_max                                     */
    const double *ljene;       /* [ntype*ntype] epsilon matrix                   */
    const double *ljlensq;     /* [ntype*ntype] sigma^2 matrix                   */
} orc_topo;

typedef struct {
    double elecut, lwljcut, upljcut, screen;
    int ljswitch;
    int splodr;
    int ms[3];
    int cltype;                /* 0 = bare Coulomb, 2 = PME, 3 = PPPM (erdst/engmain.F90:359) */
    int boxshp;                /* 1 = periodic, 0 = open (bare Coulomb only)      */
} orc_params;

/* cell normalisation result (erdst/realcal.F90:22-26) */
typedef struct {
    double cell_normal[9];
    double cell_len[3];
    double invcell[3];
    int is_cuboid;
    double rot[9];             /* Q^T applied to coordinates, (i,j) at i+3*j */
    int open;                  /* boxshp 0: no image, Coulomb without cutoff */
} orc_box;

/* erdst/engproc.F90:1049-1070 */
void orc_inverted_cell(const double *cell, double *invcl, double *volume);
/* erdst/realcal.F90:776-896; returns 0, or -1 if a cutoff exceeds half a box axis */
int orc_normalize_periodic(const double *cell, double elecut, double upljcut,
                           int n, const double *pos, double *pos_normal, orc_box *box);
/* one solute-solvent site pair, erdst/realcal.F90:270-355 (periodic branch) */
double orc_pair_energy(const orc_params *p, const orc_box *box,
                       const double *xa, const double *xb,
                       double qa, double qb, double eps, double sgm2);
/* erdst/realcal.F90:207-364 (refs) ; uv[nins][nmol] += real-space energy */
void orc_realcal_refs(const orc_topo *t, const orc_params *p, const orc_box *box,
                      const double *pos_normal, int nslt, const double *slt_charge,
                      const int *slt_ljtype, int nins, const double *slt_pos_normal,
                      double *uv);
/* erdst/realcal.F90:51-205 (soln): solute is molecule tagslt of the system */
void orc_realcal_soln(const orc_topo *t, const orc_params *p, const orc_box *box,
                      const double *pos_normal, int tagslt, double *uv_row);
/* erdst/realcal.F90:669-771 */
double orc_realcal_self(const orc_params *p, const orc_box *box, int nslt,
                        const double *slt_charge, const double *slt_pos_normal);

/* erdst/spline.F90:25-58 */
double orc_spline_value(int order, double x);
/* erdst/recpcal.F90:97-119 */
void orc_init_spline_axis(int order, int K, double *splfc);
/* erdst/recpcal.F90:121-163 ; engfac[(K1/2+1)*K2*K3] */
void orc_engfac(const orc_params *p, const double *invcl, double volume, double *engfac);
/* PPPM optimal influence function, erdst/recpcal.F90:179-337 (same layout);
 * returns 0, or -1 for an odd order or a negative alias range */
int orc_pppm_engfac(const orc_params *p, const double *cell, const double *invcl, double volume,
                    double *engfac);
/* gf_b coefficients (recpcal.F90:179-199) and gf_denom (:201-214) */
void orc_pppm_gfb(int splodr, double *gf_b);
double orc_pppm_denom(int splodr, const double *gf_b, double x, double y, double z);
/* erdst/recpcal.F90:504-538 for one atom */
void orc_spline_atom(const orc_params *p, const double *invcl, const double *x,
                     double *w /*[3][order]*/, int *grid /*[3]*/);

/* unnormalised 3-D DFTs on a K1xK2xK3 grid, erdst/fft_iface.F90:66-98.
 * spec is (K1/2+1)*K2*K3 complex as interleaved (re,im). */
void orc_fft_r2c(const int *K, const double *in, double *spec);
void orc_fft_c2r(const int *K, const double *spec, double *out);
/* 1-D complex DFT, unnormalised, sign=-1 forward, +1 backward (in place) */
void orc_fft1d(int n, int sign, double *data /*2n*/);

/* PME chain for one insertion (erdst/recpcal.F90:429-502) -> potential grid phi[G]
 * and the half-plane self energy (erdst/recpcal.F90:482-491). */
double orc_pme_solute_grid(const orc_params *p, const double *invcl, const double *engfac,
                           int nslt, const double *slt_charge, const double *slt_pos,
                           double *phi /*[G]*/, double *spec_work /*2*Gc*/, double *rho_work /*G*/);
/* splines of n atoms: w[n][3][order], grid[n][3] (erdst/recpcal.F90:340-357) */
void orc_spline_all(const orc_params *p, const double *invcl, int n, const double *pos,
                    double *w, int *grid);
/* gather (erdst/recpcal.F90:651-710) for all molecules against one grid */
void orc_pme_gather(const orc_topo *t, const orc_params *p, const double *splw,
                    const int *splg, const double *phi, int skip_mol, double *uv_row);

/* erdst/engproc.F90:1078-1108 ; returns ret in [0, n] */
int orc_binsearch(const double *coord, int n, double v);

/* Energy-coordinate mesh: erdst/engproc.F90:171-226 (&hist -> mesh).
 * Returns pemax, writes pemax values into mesh (size >= 10000) and pesoft. */
int orc_build_mesh(double ecdmin, double ecfmns, double ecdcen, double eccore, double ecdmax,
                   double eclbin, double ecfbin, double ec0bin, double finfac, int pecore,
                   double *mesh, int *pesoft, int *err);

/* Distribution state (erdst/engmain.F90:334-348) */
typedef struct {
    int slttype, wgtslf, estype, selfcal, corrcal;
    int numslv, ermax, esmax;
    const int *uvmax, *uvsoft;   /* [numslv] */
    const double *uvcrd;         /* [ermax]  */
    const double *escrd;         /* [esmax]  */
    double temp;                 /* kT in kcal/mol */
    double volume, screen, total_charge; /* for volcorrect */
    /* accumulators */
    double *edens, *ecorr, *eself;
    double *minuv, *maxuv;       /* [numslv+1] */
    double *slnuv, *flceng;      /* [numslv] (soln) */
    double engnorm, engsmpl, avslf, voffset;
    int voffset_initialized;
    int *insdst;                 /* [ermax] scratch */
    /* first error */
    char errcode[8];
    int err_species;
    double err_energy;
    int warn_mbin;
    int cltype;                  /* volcorrect's Ewald factor for cltype 1-3 */
} orc_dist;

/* erdst/engproc.F90:321-340 */
void orc_engclear(orc_dist *d);
/* erdst/engproc.F90:858-964 ; species[nmol] uvspec (0 = skip), uv row of one insertion.
 * Returns 0, or nonzero on a halt code (errcode filled). */
int orc_update_histogram(orc_dist *d, int nmol, const int *species, const double *uv_row,
                         double uvengy0, double stat_weight_system, double stat_weight_solute);
/* erdst/engproc.F90:1022-1038 */
double orc_volcorrect(const orc_dist *d, double weight);

/* Whole frame, refs mode (erdst/engproc.F90:782-856 energies only):
 * uv[nins][nmol], uv0[nins].  pos is the raw frame, slt_pos the raw insertions
 * [nins][nslt][3].  rigid != 0 reproduces the self-energy reuse of
 * erdst/engproc.F90:839-851 for slttype=2. */
int orc_energies_refs(const orc_topo *t, const orc_params *p, const double *cell,
                      const double *pos, int nslt, const double *slt_charge,
                      const int *slt_ljtype, int nins, const double *slt_pos,
                      int rigid, double *uv, double *uv0, int nthreads);
/* soln mode for the solute molecules listed in tagslt[nslt_mol]:
 * uv[nslt_mol][nmol] (entry of the solute itself is 0), uv0[nslt_mol] */
int orc_energies_soln(const orc_topo *t, const orc_params *p, const double *cell,
                      const double *pos, int nslt_mol, const int *tagslt,
                      double *uv, double *uv0, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
