/*
 * er_oracle.c -- FP64 CPU restatement of the ERmod erdst per-frame hot path.
 *
 * TEST INFRASTRUCTURE.  This file is the checker for the HIP implementation
 * and the host-core baseline timed beside it.  Only tests/, the smoke check
 * and the benchmark's cpu_baseline leg load it.  It restates the reference
 * algorithm *as written* (brute-force real-space pairs, one PME grid per
 * insertion, binary-search histogram, outer-product ecorr) in double
 * precision, i.e. the semantics of the reference's --enable-double build.
 * Every function cites the reference lines it follows.
 */
#include "er_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* cell helpers                                                              */
/* ------------------------------------------------------------------------ */
#define C(m, i, j) ((m)[(i) + 3 * (j)])

/* erdst/engproc.F90:1049-1070 */
void orc_inverted_cell(const double *cell, double *invcl, double *volume)
{
    double v = C(cell,0,0) * C(cell,1,1) * C(cell,2,2)
             + C(cell,0,1) * C(cell,1,2) * C(cell,2,0)
             + C(cell,0,2) * C(cell,1,0) * C(cell,2,1)
             - C(cell,0,2) * C(cell,1,1) * C(cell,2,0)
             - C(cell,0,1) * C(cell,1,0) * C(cell,2,2)
             - C(cell,0,0) * C(cell,1,2) * C(cell,2,1);
    C(invcl,0,0) = C(cell,1,1) * C(cell,2,2) - C(cell,1,2) * C(cell,2,1);
    C(invcl,0,1) = C(cell,0,2) * C(cell,2,1) - C(cell,0,1) * C(cell,2,2);
    C(invcl,0,2) = C(cell,0,1) * C(cell,1,2) - C(cell,0,2) * C(cell,1,1);
    C(invcl,1,0) = C(cell,1,2) * C(cell,2,0) - C(cell,1,0) * C(cell,2,2);
    C(invcl,1,1) = C(cell,0,0) * C(cell,2,2) - C(cell,0,2) * C(cell,2,0);
    C(invcl,1,2) = C(cell,0,2) * C(cell,1,0) - C(cell,0,0) * C(cell,1,2);
    C(invcl,2,0) = C(cell,1,0) * C(cell,2,1) - C(cell,1,1) * C(cell,2,0);
    C(invcl,2,1) = C(cell,0,1) * C(cell,2,0) - C(cell,0,0) * C(cell,2,1);
    C(invcl,2,2) = C(cell,0,0) * C(cell,1,1) - C(cell,0,1) * C(cell,1,0);
    for (int k = 0; k < 9; k++) invcl[k] /= v;
    *volume = v;
}

/* Fortran anint: nearest integer, halves away from zero */
static double orc_anint(double t)
{
    return (t >= 0.0) ? floor(t + 0.5) : -floor(-t + 0.5);
}

/* erdst/realcal.F90:776-896.  Only the upper-triangular branch of
 * make_cell_uppertriangular (realcal.F90:784-790) is restated: every cell the
 * tests build is upper triangular, and the general branch divides every
 * column by |a1| (realcal.F90:799), which is not a rotation. */
int orc_normalize_periodic(const double *cell, double elecut, double upljcut,
                           int n, const double *pos, double *pos_normal, orc_box *box)
{
    memcpy(box->cell_normal, cell, 9 * sizeof(double));
    box->open = 0;
    memset(box->rot, 0, sizeof(box->rot));
    box->rot[0] = box->rot[4] = box->rot[8] = 1.0;
    const int rotate = C(cell,1,0) != 0.0 || C(cell,2,0) != 0.0 || C(cell,2,1) != 0.0;
    if (rotate) {
        /* make_cell_uppertriangular (realcal.F90:776-817): cell = QR by
         * modified Gram-Schmidt, cell_normal = R, positions Q^T x.  Each
         * column is normalised by its own length; the reference takes the
         * length of column 1 for all three (line 799), which makes Q no
         * rotation. */
        double t[9], q[9], r[9] = {0};
        memcpy(t, cell, sizeof t);
        for (int i = 0; i < 3; i++) {
            double len = sqrt(C(t,0,i) * C(t,0,i) + C(t,1,i) * C(t,1,i) + C(t,2,i) * C(t,2,i));
            for (int k = 0; k < 3; k++) C(q,k,i) = C(t,k,i) / len;
            C(r,i,i) = len;
            for (int j = i + 1; j < 3; j++) {
                double d = C(q,0,i) * C(t,0,j) + C(q,1,i) * C(t,1,j) + C(q,2,i) * C(t,2,j);
                C(r,i,j) = d;
                for (int k = 0; k < 3; k++) C(t,k,j) -= d * C(q,k,i);
            }
        }
        memcpy(box->cell_normal, r, sizeof r);
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) C(box->rot,i,k) = C(q,k,i);   /* rot = Q^T */
    }
    /* normalize_cell_vector, realcal.F90:826-830: the test re-reads (1,1)
     * after each flip, so only the first column is ever negated */
    for (int i = 0; i < 3; i++)
        if (C(box->cell_normal, 0, 0) < 0.0)
            for (int r = 0; r < 3; r++) C(box->cell_normal, r, i) = -C(box->cell_normal, r, i);
    /* its lattice reduction, realcal.F90:832-847: columns 2 and 3 shifted by
     * whole multiples of columns 1 and 2 when an off-diagonal element
     * exceeds half the diagonal by more than cutoff_thres = 1e-4.  The
     * multiple is anint(element / |diagonal|) with this frame's diagonal
     * (the reference reads invcell_normal as its previous frame left it,
     * the same value for a constant cell from the second frame on). */
    {
        double *cn = box->cell_normal;
        const double thres = 1.0e-4;
        if (fabs(C(cn,0,1)) > 0.5 * C(cn,0,0) + thres) {
            const double f = orc_anint(C(cn,0,1) / fabs(C(cn,0,0)));
            for (int r = 0; r < 3; r++) C(cn,r,1) -= C(cn,r,0) * f;
        }
        if (fabs(C(cn,0,2)) > 0.5 * C(cn,0,0) + thres) {
            const double f = orc_anint(C(cn,0,2) / fabs(C(cn,0,0)));
            for (int r = 0; r < 3; r++) C(cn,r,2) -= C(cn,r,0) * f;
        }
        if (fabs(C(cn,1,2)) > 0.5 * C(cn,1,1) + thres) {
            const double f = orc_anint(C(cn,1,2) / fabs(C(cn,1,1)));
            for (int r = 0; r < 3; r++) C(cn,r,2) -= C(cn,r,1) * f;
        }
    }
    /* realcal.F90:849-856 */
    double cutoff = elecut < upljcut ? elecut : upljcut;
    for (int i = 0; i < 3; i++)
        if (cutoff > C(box->cell_normal, i, i) * 0.5) return -1;
    /* realcal.F90:868-879 */
    for (int i = 0; i < 3; i++) {
        box->cell_len[i] = fabs(C(box->cell_normal, i, i));
        box->invcell[i] = 1.0 / box->cell_len[i];
    }
    box->is_cuboid = !(fabs(C(cell,0,1)) > 1e-8 || fabs(C(cell,0,2)) > 1e-8 || fabs(C(cell,1,2)) > 1e-8);
    /* realcal.F90:885-895: shift Z, then Y, then X */
    for (int a = 0; a < n; a++) {
        double x[3] = {pos[3*a], pos[3*a+1], pos[3*a+2]};
        if (rotate) {
            double y[3];
            for (int i = 0; i < 3; i++)
                y[i] = C(box->rot,i,0) * x[0] + C(box->rot,i,1) * x[1] + C(box->rot,i,2) * x[2];
            x[0] = y[0]; x[1] = y[1]; x[2] = y[2];
        }
        for (int ax = 2; ax >= 0; ax--) {
            double f = floor(box->invcell[ax] * x[ax]);
            for (int r = 0; r < 3; r++) x[r] -= C(box->cell_normal, r, ax) * f;
        }
        pos_normal[3*a] = x[0]; pos_normal[3*a+1] = x[1]; pos_normal[3*a+2] = x[2];
    }
    return 0;
}

/* minimum image, realcal.F90:271-284 */
static void min_image(const orc_box *box, double *d)
{
    if (box->open) return;   /* boxshp 0: no periodic images */
    if (box->is_cuboid) {
        for (int k = 0; k < 3; k++) {
            double h = 0.5 * box->cell_len[k];
            d[k] = h - fabs(h - fabs(d[k]));
        }
    } else {
        for (int ax = 2; ax >= 0; ax--) {
            /* Fortran anint rounds half away from zero */
            double t = d[ax] * box->invcell[ax];
            double f = (t >= 0.0) ? floor(t + 0.5) : -floor(-t + 0.5);
            for (int r = 0; r < 3; r++) d[r] -= C(box->cell_normal, r, ax) * f;
        }
    }
}

/* realcal.F90:899-915 */
static void gmx_force_params(int pw, double lw, double up, double *A, double *B, double *Cc)
{
    double df = up - lw;
    *A = -(double)pw * ((double)(pw + 4) * up - (double)(pw + 1) * lw)
         / (pow(up, pw + 2) * df * df) / 3.0;
    *B = (double)pw * ((double)(pw + 3) * up - (double)(pw + 1) * lw)
         / (pow(up, pw + 2) * df * df * df) / 4.0;
    *Cc = 1.0 / pow(up, pw) - (*A) * df * df * df - (*B) * df * df * df * df;
}

/* realcal.F90:285-355 for one pair at squared distance dis2 */
static double pair_from_dis2(const orc_params *p, double dis2, double qa, double qb,
                             double ljeps, double ljsgm2)
{
    double rst = sqrt(dis2);
    double eplj = 0.0, epcl = 0.0;
    if (!(rst > p->upljcut)) {
        double invr2 = ljsgm2 / dis2;
        double invr6 = invr2 * invr2 * invr2;
        double lw = p->lwljcut, up = p->upljcut;
        switch (p->ljswitch) {
        case ORC_LJSWT_POT_CHM:
        case ORC_LJSWT_POT_GMX:
            eplj = 4.0 * ljeps * invr6 * (invr6 - 1.0);
            if (rst > lw) {
                double swth;
                if (p->ljswitch == ORC_LJSWT_POT_CHM) {
                    double lw2 = lw * lw, up2 = up * up;
                    double t = dis2 - up2, den = up2 - lw2;
                    swth = (2.0 * dis2 + up2 - 3.0 * lw2) * (t * t) / (den * den * den);
                } else {
                    double s = (rst - lw) / (up - lw);
                    swth = 1.0 - 10.0 * s * s * s + 15.0 * s * s * s * s - 6.0 * s * s * s * s * s;
                }
                eplj = swth * eplj;
            }
            break;
        case ORC_LJSWT_FRC_CHM: {
            double lw3 = lw * lw * lw, up3 = up * up * up, lw6 = lw3 * lw3, up6 = up3 * up3;
            double sg6 = ljsgm2 * ljsgm2 * ljsgm2, vdwa, vdwb;
            if (rst <= lw) {
                vdwa = invr6 * invr6 - sg6 * sg6 / (lw6 * up6);
                vdwb = invr6 - sg6 / (lw3 * up3);
            } else {
                double invr3 = sqrt(invr6), sg3 = sqrt(sg6);
                double ta = invr6 - sg6 / up6, tb = invr3 - sg3 / up3;
                vdwa = up6 / (up6 - lw6) * ta * ta;
                vdwb = up3 / (up3 - lw3) * tb * tb;
            }
            eplj = 4.0 * ljeps * (vdwa - vdwb);
            break;
        }
        case ORC_LJSWT_FRC_GMX: {
            double rA, rB, rC, aA, aB, aC;
            gmx_force_params(12, lw, up, &rA, &rB, &rC);
            gmx_force_params(6, lw, up, &aA, &aB, &aC);
            double sg6 = ljsgm2 * ljsgm2 * ljsgm2, vdwa, vdwb;
            if (rst <= lw) {
                vdwa = invr6 * invr6 - sg6 * sg6 * rC;
                vdwb = invr6 - sg6 * aC;
            } else {
                double s = rst - lw, s3 = s * s * s, s4 = s3 * s;
                vdwa = invr6 * invr6 - sg6 * sg6 * (rA * s3 + rB * s4 + rC);
                vdwb = invr6 - sg6 * (aA * s3 + aB * s4 + aC);
            }
            eplj = 4.0 * ljeps * (vdwa - vdwb);
            break;
        }
        default:
            eplj = NAN;
        }
    }
    if (!(rst >= p->elecut))   /* bare Coulomb for cltype 0 (realcal.F90:502-506) */
        epcl = (p->cltype == 0) ? qa * qb / rst : qa * qb * erfc(p->screen * rst) / rst;
    return eplj + epcl;
}

double orc_pair_energy(const orc_params *p, const orc_box *box,
                       const double *xa, const double *xb,
                       double qa, double qb, double eps, double sgm2)
{
    double d[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
    min_image(box, d);
    double dis2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (box->open) {          /* reelcut = infinity for an open system (realcal.F90:388) */
        orc_params q = *p;
        q.elecut = INFINITY;
        return pair_from_dis2(&q, dis2, qa, qb, eps, sgm2);
    }
    return pair_from_dis2(p, dis2, qa, qb, eps, sgm2);
}

/* erdst/realcal.F90:207-364 -- brute force over every
 * (insertion, molecule, solute site, solvent site) quadruple */
void orc_realcal_refs(const orc_topo *t, const orc_params *p, const orc_box *box,
                      const double *pos_normal, int nslt, const double *slt_charge,
                      const int *slt_ljtype, int nins, const double *slt_pos_normal,
                      double *uv)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int cnt = 0; cnt < nins; cnt++) {
        for (int i = 0; i < t->nmol; i++) {
            double e = 0.0;
            for (int is = 0; is < nslt; is++) {
                const double *xa = slt_pos_normal + 3 * ((size_t)cnt * nslt + is);
                for (int atj = t->mol_begin[i]; atj < t->mol_begin[i + 1]; atj++) {
                    int ti = slt_ljtype[is], tj = t->ljtype[atj];
                    e += orc_pair_energy(p, box, xa, pos_normal + 3 * atj, slt_charge[is],
                                         t->charge[atj], t->ljene[ti + t->ntype * tj],
                                         t->ljlensq[ti + t->ntype * tj]);
                }
            }
            uv[(size_t)cnt * t->nmol + i] += e;
        }
    }
}

/* erdst/realcal.F90:51-205 */
void orc_realcal_soln(const orc_topo *t, const orc_params *p, const orc_box *box,
                      const double *pos_normal, int tagslt, double *uv_row)
{
    for (int i = 0; i < t->nmol; i++) {
        if (i == tagslt) continue;
        double e = 0.0;
        for (int ati = t->mol_begin[tagslt]; ati < t->mol_begin[tagslt + 1]; ati++)
            for (int atj = t->mol_begin[i]; atj < t->mol_begin[i + 1]; atj++) {
                int ti = t->ljtype[ati], tj = t->ljtype[atj];
                e += orc_pair_energy(p, box, pos_normal + 3 * ati, pos_normal + 3 * atj,
                                     t->charge[ati], t->charge[atj],
                                     t->ljene[ti + t->ntype * tj], t->ljlensq[ti + t->ntype * tj]);
            }
        uv_row[i] += e;
    }
}

/* erdst/realcal.F90:669-771 */
double orc_realcal_self(const orc_params *p, const orc_box *box, int nslt,
                        const double *q, const double *x)
{
    double e = 0.0;
    for (int is = 0; is < nslt; is++) {
        e += -q[is] * q[is] * p->screen / sqrt(ORC_PI);
        for (int js = is + 1; js < nslt; js++) {
            double d[3] = {x[3*is] - x[3*js], x[3*is+1] - x[3*js+1], x[3*is+2] - x[3*js+2]};
            min_image(box, d);
            double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            e += -q[is] * q[js] * erf(p->screen * r) / r;
        }
    }
    return e;
}

/* ------------------------------------------------------------------------ */
/* PME                                                                       */
/* ------------------------------------------------------------------------ */

/* erdst/spline.F90:25-58: M_n(x) = sum_{i<=int(x)} coeff(i) (x-i)^(n-1),
 * coeff(i) = (-1)^i n / (i! (n-i)!) */
double orc_spline_value(int order, double x)
{
    double f = 0.0;
    if (x > 0.0 && x < (double)order) {
        int k = (int)x;
        for (int i = 0; i <= k; i++) {
            double fac = 1.0;
            for (int m = 1; m <= i; m++) fac *= (double)m;
            for (int m = 1; m <= order - i; m++) fac *= (double)m;
            double c = (double)order / fac;
            if (i % 2 == 1) c = -c;
            f += c * pow(x - (double)i, order - 1);
        }
    }
    return f;
}

/* erdst/recpcal.F90:97-119 */
void orc_init_spline_axis(int order, int K, double *splfc)
{
    for (int rci = 0; rci < K; rci++) {
        double re = 0.0, im = 0.0;
        for (int spi = 0; spi <= order - 2; spi++) {
            double chr = orc_spline_value(order, (double)(spi + 1));
            double a = 2.0 * ORC_PI * (double)(spi * rci) / (double)K;
            re += chr * cos(a);
            im += chr * sin(a);
        }
        splfc[rci] = re * re + im * im;
    }
}

/* erdst/recpcal.F90:121-163 */
void orc_engfac(const orc_params *p, const double *invcl, double volume, double *engfac)
{
    const int K1 = p->ms[0], K2 = p->ms[1], K3 = p->ms[2], H = K1 / 2 + 1;
    double *s1 = malloc(sizeof(double) * K1), *s2 = malloc(sizeof(double) * K2),
           *s3 = malloc(sizeof(double) * K3);
    orc_init_spline_axis(p->splodr, K1, s1);
    orc_init_spline_axis(p->splodr, K2, s2);
    orc_init_spline_axis(p->splodr, K3, s3);
    for (int r3 = 0; r3 < K3; r3++)
        for (int r2 = 0; r2 < K2; r2++)
            for (int r1 = 0; r1 < H; r1++) {
                double factor = 0.0;
                int rc[3] = {r1, r2, r3}, Km[3] = {K1, K2, K3};
                double inm[3];
                int skip = 0;
                if (r1 == 0 && r2 == 0 && r3 == 0) skip = 1;
                for (int m = 0; m < 3 && !skip; m++) {
                    if ((p->splodr % 2 == 1) && (2 * abs(rc[m]) == Km[m])) skip = 1;
                    inm[m] = (rc[m] <= Km[m] / 2) ? (double)rc[m] : (double)(rc[m] - Km[m]);
                }
                if (!skip) {
                    double xs[3], r2s = 0.0;
                    for (int m = 0; m < 3; m++) {
                        /* dot_product(invcl(:, m), inm(:)) */
                        xs[m] = C(invcl,0,m) * inm[0] + C(invcl,1,m) * inm[1] + C(invcl,2,m) * inm[2];
                        r2s += xs[m] * xs[m];
                    }
                    double chr = ORC_PI * ORC_PI * r2s / (p->screen * p->screen);
                    factor = exp(-chr) / r2s / ORC_PI / volume;
                    factor = factor / (s1[r1] * s2[r2] * s3[r3]);
                }
                engfac[r1 + H * (r2 + K2 * r3)] = factor;
            }
    engfac[0] = 0.0;
    free(s1); free(s2); free(s3);
}

/* calc_gfb_pppm, erdst/recpcal.F90:179-199: coefficients of the aliasing-sum
 * polynomial in sin^2; the final update of gf_b(1) uses the inner loop's
 * exit value of l, which is 0 */
void orc_pppm_gfb(int splodr, double *gf_b)
{
    for (int l = 1; l < splodr; l++) gf_b[l] = 0.0;
    gf_b[0] = 1.0;
    for (int m = 1; m <= splodr - 1; m++) {
        for (int l = m; l >= 1; l--)
            gf_b[l] = 4.0 * (gf_b[l] * (l - m) * (l - m - 0.5) - gf_b[l - 1] * (l - m - 1) * (l - m - 1));
        gf_b[0] = 4.0 * (gf_b[0] * (0 - m) * (0 - m - 0.5));
    }
    double fact = 1.0;
    for (int n = 2; n <= 2 * splodr - 1; n++) fact *= n;
    for (int m = 0; m < splodr; m++) gf_b[m] /= fact;
}

/* gf_denom, erdst/recpcal.F90:201-214 */
double orc_pppm_denom(int splodr, const double *gf_b, double x, double y, double z)
{
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int l = splodr - 1; l >= 0; l--) {
        s1 = gf_b[l] + s1 * x;
        s2 = gf_b[l] + s2 * y;
        s3 = gf_b[l] + s3 * z;
    }
    double pr = s1 * s2 * s3;
    return pr * pr;
}

/* recpcal_pppm_greenfunc, erdst/recpcal.F90:217-337: for each wave vector k,
 * 4 pi / V * sum over aliases km = k + 2 pi invcl^T (mx K1, my K2, mz K2) of
 * (k.km / km^2) exp(-km^2 / 4 screen^2) U(km)^2, over k^2 gf_denom; the
 * alias range from LAMMPS' bound with EPS_HOC = 1e-7.  The third alias
 * component is mz*K2 as in the reference (equal to mz*K3 on the grids where
 * K2 = K3). */
int orc_pppm_engfac(const orc_params *p, const double *cell, const double *invcl, double volume,
                    double *engfac)
{
    const int K1 = p->ms[0], K2 = p->ms[1], K3 = p->ms[2], H = K1 / 2 + 1, P = p->splodr;
    if (P % 2 == 1) return -1;
    double gf_b[16];
    orc_pppm_gfb(P, gf_b);
    const double root = pow(-log(1.0e-7), 0.25);
    const double mmax[3] = {p->screen / (ORC_PI * K1) * root, p->screen / (ORC_PI * K2) * root,
                            p->screen / (ORC_PI * K3) * root};
    int amax[3];
    for (int d = 0; d < 3; d++)
        amax[d] = (int)(C(cell,0,d) * mmax[0] + C(cell,1,d) * mmax[1] + C(cell,2,d) * mmax[2]);
    if (amax[0] < 0 || amax[1] < 0 || amax[2] < 0) return -1;
    const int Km[3] = {K1, K2, K3};
    for (int r3 = 0; r3 < K3; r3++)
        for (int r2 = 0; r2 < K2; r2++)
            for (int r1 = 0; r1 < H; r1++) {
                const int rc[3] = {r1, r2, r3};
                double inm[3], s2h[3], k[3], k2 = 0.0, factor = 0.0;
                for (int d = 0; d < 3; d++) {
                    inm[d] = (rc[d] <= Km[d] / 2) ? rc[d] : rc[d] - Km[d];
                    double sv = sin(ORC_PI * inm[d] / Km[d]);
                    s2h[d] = sv * sv;
                }
                for (int d = 0; d < 3; d++) {
                    k[d] = 2.0 * ORC_PI * (C(invcl,0,d) * inm[0] + C(invcl,1,d) * inm[1] + C(invcl,2,d) * inm[2]);
                    k2 += k[d] * k[d];
                }
                if (!(r1 == 0 && r2 == 0 && r3 == 0) && k2 > 0.0) {
                    double sum = 0.0;
                    for (int mx = -amax[0]; mx <= amax[0]; mx++) {
                        double t1 = ORC_PI * (inm[0] / K1 + mx);
                        double u1 = (t1 == 0.0) ? 1.0 : pow(sin(t1) / t1, 2 * P);
                        for (int my = -amax[1]; my <= amax[1]; my++) {
                            double t2 = ORC_PI * (inm[1] / K2 + my);
                            double u2 = (t2 == 0.0) ? 1.0 : pow(sin(t2) / t2, 2 * P);
                            for (int mz = -amax[2]; mz <= amax[2]; mz++) {
                                double t3 = ORC_PI * (inm[2] / K3 + mz);
                                double u3 = (t3 == 0.0) ? 1.0 : pow(sin(t3) / t3, 2 * P);
                                const double mN[3] = {(double)mx * K1, (double)my * K2, (double)mz * K2};
                                double km[3], kdk = 0.0, km2 = 0.0, g = 1.0;
                                for (int d = 0; d < 3; d++) {
                                    km[d] = k[d] + 2.0 * ORC_PI * (C(invcl,0,d) * mN[0] + C(invcl,1,d) * mN[1] +
                                                                   C(invcl,2,d) * mN[2]);
                                    g *= exp(-0.25 * (km[d] / p->screen) * (km[d] / p->screen));
                                    kdk += k[d] * km[d];
                                    km2 += km[d] * km[d];
                                }
                                sum += kdk * g * (u1 * u2 * u3) / km2;
                            }
                        }
                    }
                    factor = sum * 4.0 * ORC_PI / (k2 * orc_pppm_denom(P, gf_b, s2h[0], s2h[1], s2h[2])) / volume;
                }
                engfac[r1 + H * (r2 + K2 * r3)] = factor;
            }
    engfac[0] = 0.0;
    return 0;
}

/* the influence function of the run's cltype (engproc.F90:425-432); the
 * callers have checked a PPPM order is even */
static void influence(const orc_params *p, const double *cell, const double *invcl, double volume,
                      double *engfac)
{
    if (p->cltype == 3) (void)orc_pppm_engfac(p, cell, invcl, volume, engfac);
    else orc_engfac(p, invcl, volume, engfac);
}

/* erdst/recpcal.F90:504-538 */
void orc_spline_atom(const orc_params *p, const double *invcl, const double *x,
                     double *w, int *grid)
{
    for (int k = 0; k < 3; k++) {
        /* dot_product(invcl(k,:), xst(:)) */
        double f = C(invcl,k,0) * x[0] + C(invcl,k,1) * x[1] + C(invcl,k,2) * x[2];
        f = f - floor(f);
        double factor = f * (double)p->ms[k];
        int rci = (int)factor;
        for (int spi = 0; spi < p->splodr; spi++)
            w[k * p->splodr + spi] = orc_spline_value(p->splodr, factor - (double)(rci - spi));
        grid[k] = rci;
    }
}

static int pmod(int a, int K) { int r = a % K; return r < 0 ? r + K : r; }

/* -------- 1-D mixed-radix FFT (Cooley-Tukey, recursive, any n) ----------
 * tw[k] = exp(-2 pi i k / N) for the top-level length N; a sub-transform of
 * length n uses W_n^e = tw[e * (N/n) mod N]. */
static void fft_rec(int n, int stride, int sign, const double *in, double *out, double *tmp,
                    const double *tw, int N)
{
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    int r = 0;
    for (int f = 2; f * f <= n; f++) if (n % f == 0) { r = f; break; }
    if (r == 0) r = n;                       /* prime length: direct DFT */
    int step = N / n;
    if (r == n) {
        for (int k = 0; k < n; k++) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < n; j++) {
                int e = (int)(((long)j * k % n) * step);
                double c = tw[2 * e], s = sign * tw[2 * e + 1] * -1.0;
                const double *v = in + 2 * (size_t)j * stride;
                sr += v[0] * c - v[1] * s;
                si += v[0] * s + v[1] * c;
            }
            out[2 * k] = sr; out[2 * k + 1] = si;
        }
        return;
    }
    int m = n / r;
    for (int q = 0; q < r; q++)
        fft_rec(m, stride * r, sign, in + 2 * (size_t)q * stride, tmp + 2 * (size_t)q * m, out, tw, N);
    for (int k = 0; k < m; k++)
        for (int s = 0; s < r; s++) {
            int kk = k + s * m;
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < r; q++) {
                int e = (int)(((long)q * kk % n) * step);
                double c = tw[2 * e], sn = sign * tw[2 * e + 1] * -1.0;
                const double *v = tmp + 2 * ((size_t)q * m + k);
                sr += v[0] * c - v[1] * sn;
                si += v[0] * sn + v[1] * c;
            }
            out[2 * kk] = sr; out[2 * kk + 1] = si;
        }
}

static double *make_tw(int N)
{
    double *tw = malloc(sizeof(double) * 2 * N);
    for (int k = 0; k < N; k++) {
        double a = -2.0 * ORC_PI * (double)k / (double)N;
        tw[2 * k] = cos(a); tw[2 * k + 1] = sin(a);
    }
    return tw;
}

void orc_fft1d(int n, int sign, double *data)
{
    double *out = malloc(sizeof(double) * 2 * n), *tmp = malloc(sizeof(double) * 2 * n);
    double *tw = make_tw(n);
    fft_rec(n, 1, sign, data, out, tmp, tw, n);
    memcpy(data, out, sizeof(double) * 2 * n);
    free(out); free(tmp); free(tw);
}

/* full complex 3-D transform in place, grid index x1 + K1*(x2 + K2*x3) */
static void fft3d_complex(const int *K, int sign, double *g)
{
    int K1 = K[0], K2 = K[1], K3 = K[2], nmax = K1 > K2 ? K1 : K2;
    if (K3 > nmax) nmax = K3;
    double *line = malloc(sizeof(double) * 2 * nmax), *out = malloc(sizeof(double) * 2 * nmax),
           *tmp = malloc(sizeof(double) * 2 * nmax);
    double *tw1 = make_tw(K1), *tw2 = make_tw(K2), *tw3 = make_tw(K3);
    for (int x3 = 0; x3 < K3; x3++)
        for (int x2 = 0; x2 < K2; x2++) {
            double *p = g + 2 * ((size_t)K1 * (x2 + (size_t)K2 * x3));
            fft_rec(K1, 1, sign, p, out, tmp, tw1, K1);
            memcpy(p, out, sizeof(double) * 2 * K1);
        }
    for (int x3 = 0; x3 < K3; x3++)
        for (int x1 = 0; x1 < K1; x1++) {
            for (int x2 = 0; x2 < K2; x2++) {
                size_t o = 2 * (x1 + (size_t)K1 * (x2 + (size_t)K2 * x3));
                line[2 * x2] = g[o]; line[2 * x2 + 1] = g[o + 1];
            }
            fft_rec(K2, 1, sign, line, out, tmp, tw2, K2);
            for (int x2 = 0; x2 < K2; x2++) {
                size_t o = 2 * (x1 + (size_t)K1 * (x2 + (size_t)K2 * x3));
                g[o] = out[2 * x2]; g[o + 1] = out[2 * x2 + 1];
            }
        }
    for (int x2 = 0; x2 < K2; x2++)
        for (int x1 = 0; x1 < K1; x1++) {
            for (int x3 = 0; x3 < K3; x3++) {
                size_t o = 2 * (x1 + (size_t)K1 * (x2 + (size_t)K2 * x3));
                line[2 * x3] = g[o]; line[2 * x3 + 1] = g[o + 1];
            }
            fft_rec(K3, 1, sign, line, out, tmp, tw3, K3);
            for (int x3 = 0; x3 < K3; x3++) {
                size_t o = 2 * (x1 + (size_t)K1 * (x2 + (size_t)K2 * x3));
                g[o] = out[2 * x3]; g[o + 1] = out[2 * x3 + 1];
            }
        }
    free(line); free(out); free(tmp); free(tw1); free(tw2); free(tw3);
}

/* erdst/fft_iface.F90:66-82 -- forward (e^{-i}), unnormalised, half spectrum
 * along the first axis */
void orc_fft_r2c(const int *K, const double *in, double *spec)
{
    size_t G = (size_t)K[0] * K[1] * K[2];
    int H = K[0] / 2 + 1;
    double *g = malloc(sizeof(double) * 2 * G);
    for (size_t i = 0; i < G; i++) { g[2 * i] = in[i]; g[2 * i + 1] = 0.0; }
    fft3d_complex(K, -1, g);
    for (int x3 = 0; x3 < K[2]; x3++)
        for (int x2 = 0; x2 < K[1]; x2++)
            for (int m1 = 0; m1 < H; m1++) {
                size_t s = m1 + (size_t)H * (x2 + (size_t)K[1] * x3);
                size_t o = m1 + (size_t)K[0] * (x2 + (size_t)K[1] * x3);
                spec[2 * s] = g[2 * o]; spec[2 * s + 1] = g[2 * o + 1];
            }
    free(g);
}

/* erdst/fft_iface.F90:84-98 -- backward (e^{+i}), unnormalised; the missing
 * half is the Hermitian mirror X(-m) = conj X(m) */
void orc_fft_c2r(const int *K, const double *spec, double *out)
{
    int K1 = K[0], K2 = K[1], K3 = K[2], H = K1 / 2 + 1;
    size_t G = (size_t)K1 * K2 * K3;
    double *g = malloc(sizeof(double) * 2 * G);
    for (int m3 = 0; m3 < K3; m3++)
        for (int m2 = 0; m2 < K2; m2++)
            for (int m1 = 0; m1 < K1; m1++) {
                size_t o = m1 + (size_t)K1 * (m2 + (size_t)K2 * m3);
                if (m1 < H) {
                    size_t s = m1 + (size_t)H * (m2 + (size_t)K2 * m3);
                    g[2 * o] = spec[2 * s]; g[2 * o + 1] = spec[2 * s + 1];
                } else {
                    int n1 = K1 - m1, n2 = (K2 - m2) % K2, n3 = (K3 - m3) % K3;
                    size_t s = n1 + (size_t)H * (n2 + (size_t)K2 * n3);
                    g[2 * o] = spec[2 * s]; g[2 * o + 1] = -spec[2 * s + 1];
                }
            }
    fft3d_complex(K, +1, g);
    for (size_t i = 0; i < G; i++) out[i] = g[2 * i];
    free(g);
}

/* erdst/recpcal.F90:429-502 for one insertion (and :359-427 for soln) */
double orc_pme_solute_grid(const orc_params *p, const double *invcl, const double *engfac,
                           int nslt, const double *q, const double *x,
                           double *phi, double *spec, double *rho)
{
    const int K1 = p->ms[0], K2 = p->ms[1], K3 = p->ms[2], n = p->splodr, H = K1 / 2 + 1;
    size_t G = (size_t)K1 * K2 * K3;
    memset(rho, 0, sizeof(double) * G);
    double w[3 * 16];
    int grid[3];
    for (int s = 0; s < nslt; s++) {
        orc_spline_atom(p, invcl, x + 3 * s, w, grid);
        for (int c3 = 0; c3 < n; c3++)
            for (int c2 = 0; c2 < n; c2++)
                for (int c1 = 0; c1 < n; c1++) {
                    int r1 = pmod(grid[0] - c1, K1), r2 = pmod(grid[1] - c2, K2), r3 = pmod(grid[2] - c3, K3);
                    rho[r1 + (size_t)K1 * (r2 + (size_t)K2 * r3)] +=
                        q[s] * w[c1] * w[n + c2] * w[2 * n + c3];
                }
    }
    orc_fft_r2c(p->ms, rho, spec);
    /* recpcal.F90:482-491 */
    double self = 0.0;
    for (int m3 = 0; m3 < K3; m3++)
        for (int m2 = 0; m2 < K2; m2++)
            for (int m1 = 0; m1 < H; m1++) {
                size_t s = m1 + (size_t)H * (m2 + (size_t)K2 * m3);
                double wgt;
                if (K1 % 2 == 0) wgt = (m1 == 0 || m1 == H - 1) ? 0.5 : 1.0;
                else wgt = (m1 == 0) ? 0.5 : 1.0;
                self += wgt * engfac[s] * (spec[2 * s] * spec[2 * s] + spec[2 * s + 1] * spec[2 * s + 1]);
            }
    /* recpcal.F90:493-499 */
    size_t Gc = (size_t)H * K2 * K3;
    for (size_t s = 0; s < Gc; s++) { spec[2 * s] *= engfac[s]; spec[2 * s + 1] *= engfac[s]; }
    orc_fft_c2r(p->ms, spec, phi);
    return self;
}

/* erdst/recpcal.F90:340-357 + 504-538: splines of every atom of the frame */
void orc_spline_all(const orc_params *p, const double *invcl, int n, const double *pos,
                    double *w, int *grid)
{
    for (int a = 0; a < n; a++)
        orc_spline_atom(p, invcl, pos + 3 * (size_t)a, w + (size_t)a * 3 * p->splodr, grid + 3 * (size_t)a);
}

/* erdst/recpcal.F90:651-710 (refs) / :592-649 (soln), splines precomputed */
void orc_pme_gather(const orc_topo *t, const orc_params *p, const double *splw,
                    const int *splg, const double *phi, int skip_mol, double *uv_row)
{
    const int K1 = p->ms[0], K2 = p->ms[1], K3 = p->ms[2], n = p->splodr;
    for (int i = 0; i < t->nmol; i++) {
        if (i == skip_mol) continue;
        double e = 0.0;
        for (int a = t->mol_begin[i]; a < t->mol_begin[i + 1]; a++) {
            const double *w = splw + (size_t)a * 3 * n;
            const int *grid = splg + 3 * (size_t)a;
            for (int c3 = 0; c3 < n; c3++) {
                double f1 = t->charge[a] * w[2 * n + c3];
                int r3 = pmod(grid[2] - c3, K3);
                for (int c2 = 0; c2 < n; c2++) {
                    double f2 = f1 * w[n + c2];
                    int r2 = pmod(grid[1] - c2, K2);
                    const double *row = phi + (size_t)K1 * (r2 + (size_t)K2 * r3);
                    for (int c1 = 0; c1 < n; c1++)
                        e += f2 * w[c1] * row[pmod(grid[0] - c1, K1)];
                }
            }
        }
        uv_row[i] += e;
    }
}

/* ------------------------------------------------------------------------ */
/* histogram                                                                 */
/* ------------------------------------------------------------------------ */

/* erdst/engproc.F90:1078-1108 (coord is FP64, v is compared in FP64) */
int orc_binsearch(const double *coord, int n, double v)
{
    if (v < coord[0]) return 0;
    if (v > coord[n - 1]) return n;
    int rmin = 1, rmax = n + 1;
    while (rmax - rmin > 1) {
        int rmid = (rmin + rmax - 1) >> 1;
        if (v > coord[rmid - 1]) rmin = rmid + 1;
        else rmax = rmid + 1;
    }
    return rmin - 1;
}

/* erdst/engproc.F90:171-226 */
int orc_build_mesh(double ecdmin, double ecfmns, double ecdcen, double eccore, double ecdmax,
                   double eclbin, double ecfbin, double ec0bin, double finfac, int pecore,
                   double *mesh, int *pesoft_out, int *err)
{
    const int rglmax = 5, large = 10000;
    double cdrgvl[7];
    int uprgcd[7];
    *err = 0;
    double ectmvl = finfac * ecfbin;
    ecdmin -= ectmvl;
    ecfmns -= ectmvl;
    double ecmns0 = ecdcen - ectmvl;
    double ecpls0 = 2.0 * ecdcen - ecmns0;
    double ecfpls = 2.0 * ecdcen - ecfmns;
    eccore = eccore + ecfpls - ecdcen;
    cdrgvl[0] = ecdmin; cdrgvl[1] = ecfmns; cdrgvl[2] = ecmns0; cdrgvl[3] = ecpls0;
    cdrgvl[4] = ecfpls; cdrgvl[5] = eccore; cdrgvl[6] = ecdmax;
    uprgcd[0] = 1;
    for (int regn = 1; regn <= rglmax; regn++) {
        double factor = (regn == 1 || regn == 5) ? eclbin : (regn == 2 || regn == 4) ? ecfbin : ec0bin;
        double incre = cdrgvl[regn] - cdrgvl[regn - 1];
        if (incre < factor) { *err = 1; return 0; }          /* eng_ecd */
        int iduv = (int)lround(incre / factor);              /* Fortran nint */
        uprgcd[regn] = uprgcd[regn - 1] + iduv;
    }
    int pesoft = uprgcd[rglmax] - uprgcd[0];
    int pemax = pesoft + pecore;
    uprgcd[rglmax + 1] = pemax;
    if (pemax > large) { *err = 2; return 0; }               /* eng_siz */
    /* ercrd is 1-based in the reference: mesh[i-1] = ercrd(i) */
    for (int r = 0; r <= rglmax + 1; r++) mesh[uprgcd[r] - 1] = cdrgvl[r];
    int nreg = (pecore == 0) ? rglmax : rglmax + 1;
    for (int regn = 1; regn <= nreg; regn++) {
        int minrg = uprgcd[regn - 1], maxrg = uprgcd[regn];
        double incre;
        if (regn <= rglmax) incre = mesh[maxrg - 1] - mesh[minrg - 1];
        else incre = log(mesh[maxrg - 1] / mesh[minrg - 1]);
        double factor = incre / (double)(maxrg - minrg);
        double base = mesh[minrg - 1];
        for (int iduv = minrg; iduv <= maxrg - 1; iduv++) {
            double inc = factor * (double)(iduv - minrg);
            if (regn <= rglmax) mesh[iduv - 1] = base + inc;
            else mesh[iduv - 1] = base * exp(inc);
        }
    }
    *pesoft_out = pesoft;
    return pemax;
}

/* erdst/engproc.F90:321-340 */
void orc_engclear(orc_dist *d)
{
    memset(d->edens, 0, sizeof(double) * d->ermax);
    if (d->corrcal) memset(d->ecorr, 0, sizeof(double) * (size_t)d->ermax * d->ermax);
    if (d->selfcal) memset(d->eself, 0, sizeof(double) * d->esmax);
    if (d->slttype == ORC_SLT_SOLN && d->slnuv) memset(d->slnuv, 0, sizeof(double) * d->numslv);
    d->avslf = 0.0; d->engnorm = 0.0; d->engsmpl = 0.0;
}

/* erdst/engproc.F90:1022-1038 (PME branch) */
double orc_volcorrect(const orc_dist *d, double weight)
{
    weight = weight * d->volume;
    if (d->cltype == 0) return weight;   /* bare Coulomb: no Ewald factor (engproc.F90:1029) */
    double factor = ORC_PI * d->total_charge * d->total_charge / (d->screen * d->screen) / d->volume / 2.0;
    return weight * exp(factor / d->temp);
}

static void set_err(orc_dist *d, const char *code, int sp, double e)
{
    if (d->errcode[0]) return;
    strncpy(d->errcode, code, 7); d->errcode[7] = 0;
    d->err_species = sp; d->err_energy = e;
}

/* erdst/engproc.F90:1111-1171 ; returns 0-based bin or -1 on halt */
static int getiduv(orc_dist *d, int pti, double v)
{
    int idmin, idnum, idpti;
    if (pti > 0) {
        idmin = 0;
        for (int k = 0; k < pti - 1; k++) idmin += d->uvmax[k];
        idnum = d->uvmax[pti - 1];
        idpti = orc_binsearch(d->uvcrd + idmin, idnum, v);
    } else {
        idmin = 0; idnum = d->esmax;
        idpti = orc_binsearch(d->escrd, idnum, v);
    }
    if (idpti <= 0) { set_err(d, "eng_min", pti, v); return -1; }
    if (idpti > idnum) { set_err(d, "eng_bug", pti, v); return -1; }
    if (d->slttype == ORC_SLT_SOLN && pti > 0 && idpti > d->uvsoft[pti - 1]) {
        set_err(d, "eng_sft", pti, v); return -1;
    }
    if (idpti == idnum && v < 1.0e3) d->warn_mbin = 1;
    return idmin + idpti - 1;
}

/* erdst/engproc.F90:858-964 */
int orc_update_histogram(orc_dist *d, int nmol, const int *species, const double *uv_row,
                         double uvengy0, double stat_weight_system, double stat_weight_solute)
{
    double engnmfc;
    if (!d->wgtslf) engnmfc = 1.0;
    else {
        if (!d->voffset_initialized) { d->voffset = uvengy0; d->voffset_initialized = 1; }
        double factor = uvengy0 - d->voffset;
        engnmfc = (d->slttype == ORC_SLT_SOLN) ? exp(factor / d->temp) : exp(-factor / d->temp);
    }
    double total_weight = stat_weight_system * stat_weight_solute;
    if (d->estype == 2) total_weight = orc_volcorrect(d, total_weight);
    engnmfc *= total_weight;
    d->engnorm += engnmfc;
    d->engsmpl += 1.0;
    d->avslf += total_weight;
    if (d->selfcal) {
        int b = getiduv(d, 0, uvengy0);
        if (b < 0) return 1;
        d->eself[b] += engnmfc;
    }
    if (uvengy0 < d->minuv[0]) d->minuv[0] = uvengy0;
    if (uvengy0 > d->maxuv[0]) d->maxuv[0] = uvengy0;
    memset(d->insdst, 0, sizeof(int) * d->ermax);
    if (d->flceng) memset(d->flceng, 0, sizeof(double) * d->numslv);
    for (int k = 0; k < nmol; k++) {
        int pti = species[k];
        if (pti == 0) continue;              /* the solute itself (i == tagslt) */
        double v = uv_row[k];
        int b = getiduv(d, pti, v);
        if (b < 0) return 1;
        d->insdst[b] += 1;
        if (d->flceng) d->flceng[pti - 1] += v;
        if (v < d->minuv[pti]) d->minuv[pti] = v;
        if (v > d->maxuv[pti]) d->maxuv[pti] = v;
    }
    if (d->slttype == ORC_SLT_SOLN && d->slnuv)
        for (int s = 0; s < d->numslv; s++) d->slnuv[s] += d->flceng[s] * engnmfc;
    for (int b = 0; b < d->ermax; b++)
        if (d->insdst[b]) d->edens[b] += engnmfc * (double)d->insdst[b];
    if (d->corrcal)
        for (int i = 0; i < d->ermax; i++) {
            int k = d->insdst[i];
            if (!k) continue;
            for (int j = 0; j < d->ermax; j++) {
                int q = d->insdst[j];
                if (!q) continue;
                d->ecorr[(size_t)i * d->ermax + j] += engnmfc * (double)k * (double)q;
            }
        }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* whole-frame drivers                                                       */
/* ------------------------------------------------------------------------ */

/* erdst/engproc.F90:782-856 (refs), PME branch, energies only */
int orc_energies_refs(const orc_topo *t, const orc_params *p, const double *cell,
                      const double *pos, int nslt, const double *slt_charge,
                      const int *slt_ljtype, int nins, const double *slt_pos,
                      int rigid, double *uv, double *uv0, int nthreads)
{
    if (p->cltype == 3 && p->splodr % 2 == 1) return 2;   /* "interpolation order for PPPM must be even" */
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    double invcl[9], volume;
    orc_inverted_cell(cell, invcl, &volume);
    size_t nall = (size_t)t->natom + (size_t)nins * nslt;
    double *all = malloc(sizeof(double) * 3 * nall), *alln = malloc(sizeof(double) * 3 * nall);
    memcpy(all, pos, sizeof(double) * 3 * t->natom);
    memcpy(all + 3 * (size_t)t->natom, slt_pos, sizeof(double) * 3 * (size_t)nins * nslt);
    orc_box box;
    memset(&box, 0, sizeof box);
    if (p->boxshp == 0) {     /* realcal_prepare without normalize_periodic */
        if (p->cltype != 0) { free(all); free(alln); return 3; }   /* set_prs */
        memcpy(alln, all, sizeof(double) * 3 * nall);
        box.open = 1;
    } else if (orc_normalize_periodic(cell, p->elecut, p->upljcut, (int)nall, all, alln, &box)) {
        free(all); free(alln); return 1;
    }
    const double *sltn = alln + 3 * (size_t)t->natom;
    memset(uv, 0, sizeof(double) * (size_t)nins * t->nmol);
    if (p->cltype == 0) {
        /* bare Coulomb (engproc.F90:819-840): pairs only; the self energy
         * terms return at once (realcal.F90:730, engproc.F90:973), so
         * uvengy0 = 0.  Each molecule's sum goes to its own row: the
         * reference's realcal_bare_refs adds into uvengy(k, cnt) with k never
         * set (realcal.F90:660), an undefined element. */
        orc_realcal_refs(t, p, &box, alln, nslt, slt_charge, slt_ljtype, nins, sltn, uv);
        for (int cnt = 0; cnt < nins; cnt++) uv0[cnt] = 0.0;
        free(all); free(alln);
        return 0;
    }
    int K1 = p->ms[0], K2 = p->ms[1], K3 = p->ms[2], H = K1 / 2 + 1;
    size_t G = (size_t)K1 * K2 * K3, Gc = (size_t)H * K2 * K3;
    double *engfac = malloc(sizeof(double) * Gc);
    influence(p, cell, invcl, volume, engfac);
    double *selfrec = malloc(sizeof(double) * nins);
    double *splw = malloc(sizeof(double) * 3 * p->splodr * (size_t)t->natom);
    int *splg = malloc(sizeof(int) * 3 * (size_t)t->natom);
    orc_spline_all(p, invcl, t->natom, pos, splw, splg);
    /* recpcal_prepare_solute_refs + recpcal_energy_refs, one grid per insertion */
#pragma omp parallel
    {
        double *phi = malloc(sizeof(double) * G), *rho = malloc(sizeof(double) * G),
               *spec = malloc(sizeof(double) * 2 * Gc);
#pragma omp for schedule(dynamic, 1)
        for (int cnt = 0; cnt < nins; cnt++) {
            selfrec[cnt] = orc_pme_solute_grid(p, invcl, engfac, nslt, slt_charge,
                                               slt_pos + 3 * (size_t)cnt * nslt, phi, spec, rho);
            orc_pme_gather(t, p, splw, splg, phi, -1, uv + (size_t)cnt * t->nmol);
        }
        free(phi); free(rho); free(spec);
    }
    /* realcal_refs */
    orc_realcal_refs(t, p, &box, alln, nslt, slt_charge, slt_ljtype, nins, sltn, uv);
    /* residual_ene_refs, engproc.F90:1001-1020 */
    double qslt = 0.0;
    for (int s = 0; s < nslt; s++) qslt += slt_charge[s];
    for (int cnt = 0; cnt < nins; cnt++)
        for (int k = 0; k < t->nmol; k++) {
            double qk = 0.0;
            for (int a = t->mol_begin[k]; a < t->mol_begin[k + 1]; a++) qk += t->charge[a];
            uv[(size_t)cnt * t->nmol + k] -= ORC_PI * qslt * qk / p->screen / p->screen / volume;
        }
    /* engproc.F90:828-855 */
    double resself = -ORC_PI * qslt * qslt / p->screen / p->screen / volume / 2.0;
    double usreal = 0.0;
    for (int cnt = 0; cnt < nins; cnt++) {
        double pairep;
        if (rigid && cnt > 0) pairep = usreal;
        else pairep = orc_realcal_self(p, &box, nslt, slt_charge, sltn + 3 * (size_t)cnt * nslt);
        usreal = pairep;
        uv0[cnt] = selfrec[cnt] + pairep + resself;
    }
    free(all); free(alln); free(engfac); free(selfrec); free(splw); free(splg);
    return 0;
}

/* erdst/engproc.F90:728-780 (soln), PME branch */
int orc_energies_soln(const orc_topo *t, const orc_params *p, const double *cell,
                      const double *pos, int nslt_mol, const int *tagslt,
                      double *uv, double *uv0, int nthreads)
{
    if (p->cltype == 3 && p->splodr % 2 == 1) return 2;   /* "interpolation order for PPPM must be even" */
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    double invcl[9], volume;
    orc_inverted_cell(cell, invcl, &volume);
    double *posn = malloc(sizeof(double) * 3 * t->natom);
    orc_box box;
    memset(&box, 0, sizeof box);
    if (p->boxshp == 0) {
        if (p->cltype != 0) { free(posn); return 3; }
        memcpy(posn, pos, sizeof(double) * 3 * t->natom);
        box.open = 1;
    } else if (orc_normalize_periodic(cell, p->elecut, p->upljcut, t->natom, pos, posn, &box)) {
        free(posn); return 1;
    }
    if (p->cltype == 0) {
        /* bare Coulomb (engproc.F90:761-769): pairs only, uvengy0 = 0 */
        memset(uv, 0, sizeof(double) * (size_t)nslt_mol * t->nmol);
        for (int c = 0; c < nslt_mol; c++) {
            orc_realcal_soln(t, p, &box, posn, tagslt[c], uv + (size_t)c * t->nmol);
            uv0[c] = 0.0;
        }
        free(posn);
        return 0;
    }
    int K1 = p->ms[0], K2 = p->ms[1], K3 = p->ms[2], H = K1 / 2 + 1;
    size_t G = (size_t)K1 * K2 * K3, Gc = (size_t)H * K2 * K3;
    double *engfac = malloc(sizeof(double) * Gc), *phi = malloc(sizeof(double) * G),
           *rho = malloc(sizeof(double) * G), *spec = malloc(sizeof(double) * 2 * Gc);
    influence(p, cell, invcl, volume, engfac);
    double *splw = malloc(sizeof(double) * 3 * p->splodr * (size_t)t->natom);
    int *splg = malloc(sizeof(int) * 3 * (size_t)t->natom);
    orc_spline_all(p, invcl, t->natom, pos, splw, splg);
    memset(uv, 0, sizeof(double) * (size_t)nslt_mol * t->nmol);
    for (int c = 0; c < nslt_mol; c++) {
        int m = tagslt[c], b = t->mol_begin[m], ns = t->mol_begin[m + 1] - b;
        double *row = uv + (size_t)c * t->nmol;
        double self = orc_pme_solute_grid(p, invcl, engfac, ns, t->charge + b, pos + 3 * b, phi, spec, rho);
        orc_realcal_soln(t, p, &box, posn, m, row);
        orc_pme_gather(t, p, splw, splg, phi, m, row);
        double qs = 0.0;
        for (int a = b; a < b + ns; a++) qs += t->charge[a];
        for (int k = 0; k < t->nmol; k++) {
            if (k == m) continue;
            double qk = 0.0;
            for (int a = t->mol_begin[k]; a < t->mol_begin[k + 1]; a++) qk += t->charge[a];
            row[k] -= ORC_PI * qs * qk / p->screen / p->screen / volume;
        }
        double pairep = orc_realcal_self(p, &box, ns, t->charge + b, posn + 3 * b);
        uv0[c] = self + pairep - ORC_PI * qs * qs / p->screen / p->screen / volume / 2.0;
    }
    free(posn); free(engfac); free(phi); free(rho); free(spec); free(splw); free(splg);
    return 0;
}
