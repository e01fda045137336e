/*
 * ORACLE — test infrastructure only (see oracle.h). Parity unpinned: none —
 * every function below is checked by tests/test_oracle.py against the
 * reference kernels built in oracle/_ref (bitwise or <=1e-14), against the
 * committed fixtures in tests/golden/, and against closed forms.
 *
 * Written from the formulas of the reference kernels; each block cites the
 * reference lines it restates. Built with -ffp-contract=off so that every
 * product and sum is rounded exactly where the reference source rounds it.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

int orc_kind(void) { return 0; }

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------- */
/* U_Q.cl:36-125                                                              */
/* ------------------------------------------------------------------------- */

/* total enthalpy of one node as written inline at U_Q.cl:73-74 */
static double node_h(const orc_flow *f)
{
    return ((f->p / 0.4) + 0.5 * f->rho * (f->u * f->u + f->v * f->v + f->w * f->w) + f->p) / f->rho;
}

void orc_uq(int64_t n_launch, orc_vec5 *U_C, orc_vec5 *Q_C, const orc_grid *Grid_P,
            const double *det, orc_flow *FLOW, orc_roe *ROE_AVER)
{
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < n_launch; ++g) {
        const orc_grid *gp = &Grid_P[g];
        /* ROE face slots ip, jp, kp, im, jm, km -> neighbours N1, N0, N4, N3, N2, N5 (U_Q.cl:51-56) */
        const int nb[6] = {gp->N1, gp->N0, gp->N4, gp->N3, gp->N2, gp->N5};
        const orc_flow *fi = &FLOW[g];
        const double sqr_i = sqrt(fi->rho);
        for (int s = 0; s < 6; ++s) {
            const orc_flow *fn = &FLOW[nb[s]];
            const double sb = sqrt(fn->rho);
            double *r = ROE_AVER[g].f[s];           /* (u, v, w, p, h, rho); p is never written */
            r[5] = sb * sqr_i;                                             /* U_Q.cl:69 */
            r[0] = (sb * fn->u + sqr_i * fi->u) / (sb + sqr_i);            /* U_Q.cl:70 */
            r[1] = (sb * fn->v + sqr_i * fi->v) / (sb + sqr_i);
            r[2] = (sb * fn->w + sqr_i * fi->w) / (sb + sqr_i);
            r[4] = (sb * node_h(fn) + sqr_i * node_h(fi)) / (sb + sqr_i);  /* U_Q.cl:73-74 */
        }
        const double D = det[g];
        const double ke = 0.5 * fi->rho * (fi->u * fi->u + fi->v * fi->v + fi->w * fi->w);
        U_C[g].N0 = D * fi->rho;                                           /* U_Q.cl:112-116 */
        U_C[g].N1 = D * fi->rho * fi->u;
        U_C[g].N2 = D * fi->rho * fi->v;
        U_C[g].N3 = D * fi->rho * fi->w;
        U_C[g].N4 = D * ((fi->p / 0.4) + ke);
        Q_C[g].N0 = fi->rho;                                               /* U_Q.cl:118-122 */
        Q_C[g].N1 = fi->rho * fi->u;
        Q_C[g].N2 = fi->rho * fi->v;
        Q_C[g].N3 = fi->rho * fi->w;
        Q_C[g].N4 = ((fi->p / 0.4) + ke);
    }
}

/* ------------------------------------------------------------------------- */
/* Neighbour walks (Viscous.cl:99-118, Weno.cl:186-207)                       */
/* ------------------------------------------------------------------------- */

/* Along one direction: idx[0..6] = node at offsets -3..+3 (idx[3] = g). */
static void walk(const orc_grid *G, int64_t g, int dir, int64_t idx[7])
{
    idx[3] = g;
    int64_t p = g, m = g;
    for (int s = 1; s <= 3; ++s) {
        const orc_grid *gp = &G[p], *gm = &G[m];
        /* dir 0 (i, zeta): +N1 -N3; dir 1 (j, eta): +N0 -N2; dir 2 (k, xi): +N4 -N5 */
        p = dir == 0 ? gp->N1 : dir == 1 ? gp->N0 : gp->N4;
        m = dir == 0 ? gm->N3 : dir == 1 ? gm->N2 : gm->N5;
        idx[3 + s] = p;
        idx[3 - s] = m;
    }
}

static double v5(const orc_vec5 *v, int c)
{
    switch (c) {
    case 0: return v->N0;
    case 1: return v->N1;
    case 2: return v->N2;
    case 3: return v->N3;
    default: return v->N4;
    }
}

/*
 * One-direction derivative with the loc-selected stencils of Viscous.cl:285-435
 * (weno_form = 0) or Weno.cl:2732-2856 (weno_form = 1). The branches are tested
 * in source order and a later match overwrites an earlier one. With no match the
 * reference leaves the value undefined; here it stays 0.
 *   f[0..6] = values at offsets -3..+3, loc[0..6] the loc codes there.
 */
static double loc_deriv(const double f[7], const int loc[7], int weno_form)
{
    const int LI = loc[3], LP1 = loc[4], LP2 = loc[5], LM1 = loc[2], LM2 = loc[1];
    double d = 0.0;
    if (LI > 0 && LI <= 52 && LP1 == 100)
        d = (1.0 / 6.0) * (11.0 * f[3] - 18.0 * f[2] + 9.0 * f[1] - 2.0 * f[0]);
    if (LI > 0 && LI <= 52 && LM1 == 100)
        d = (1.0 / 6.0) * (-11.0 * f[3] + 18.0 * f[4] - 9.0 * f[5] + 2.0 * f[6]);
    if (LI == 0 && LP1 > 0 && LP1 <= 52 && LP2 == 100)
        d = (1.0 / 6.0) * (2.0 * f[4] + 3.0 * f[3] - 6.0 * f[2] + f[1]);
    if (weno_form ? (LI == 0 && LM1 > 0 && LM1 <= 52 && LM2 == 100)
                  : (LI == 0 && LM1 > 0 && LM1 != 0 && LM2 == 100))  /* Viscous.cl:309 lacks <=52 */
        d = (1.0 / 6.0) * ((-2.0) * f[2] - 3.0 * f[3] + 6.0 * f[4] - f[5]);
    if (LI == 0 && LM2 != 100 && LM1 != 100 && LP2 != 100 && LP1 != 100)
        d = (1.0 / 12.0) * (8.0 * (f[4] - f[2]) - (f[5] - f[1]));
    if (!weno_form && LI > 0 && LM2 != 100 && LM1 != 100 && LP2 != 100 && LP1 != 100)
        d = (1.0 / 12.0) * (8.0 * (f[4] - f[2]) - (f[5] - f[1]));     /* Viscous.cl:325-332 only */
    return d;
}

/* ------------------------------------------------------------------------- */
/* Viscous.cl:23-643                                                          */
/* ------------------------------------------------------------------------- */

void orc_viscous(int64_t n_launch, const orc_vec5 *U_C, orc_vec5 *F_C, orc_vec5 *E_C, orc_vec5 *G_C,
                 orc_vec5 *Fv_C, orc_vec5 *Ev_C, orc_vec5 *Gv_C, const orc_grid *Grid_P,
                 const orc_metric *metric, const double *det, orc_flow *FLOW, const orc_params *prm)
{
    const double Mach = prm->mach, Reyl = prm->reyl, Pr = prm->pr, Free_t = prm->free_t;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < n_launch; ++g) {
        int64_t idx[3][7];
        int loc[3][7];
        for (int d = 0; d < 3; ++d) {
            walk(Grid_P, g, d, idx[d]);
            for (int s = 0; s < 7; ++s) loc[d][s] = Grid_P[idx[d][s]].loc;
        }
        /* dU[c][d] = d U_c / d(zeta, eta, xi)_d   (U1_zeta .. U5_xi) */
        double dU[5][3];
        for (int c = 0; c < 5; ++c)
            for (int d = 0; d < 3; ++d) {
                double f[7];
                for (int s = 0; s < 7; ++s) f[s] = v5(&U_C[idx[d][s]], c);
                dU[c][d] = loc_deriv(f, loc[d], 0);
            }
        const double DET_I = det[g];
        const double *M = metric[g].node;
        const double zeta_x = M[0], zeta_y = M[1], zeta_z = M[2];
        const double eta_x = M[3], eta_y = M[4], eta_z = M[5];
        const double xi_x = M[6], xi_y = M[7], xi_z = M[8];
        const double U0_I = U_C[g].N0, U1_I = U_C[g].N1, U2_I = U_C[g].N2, U3_I = U_C[g].N3, U4_I = U_C[g].N4;
        orc_flow fl = FLOW[g];

        /* Viscous.cl:439-453 — quotient rule; "t" holds the internal-energy gradient */
        double ug[3], vg[3], wg[3], tg[3];
        for (int d = 0; d < 3; ++d) {
            ug[d] = (1.0 / U0_I) * (dU[1][d]) - (U1_I / ((U0_I * U0_I))) * (dU[0][d]);
            vg[d] = (1.0 / U0_I) * (dU[2][d]) - (U2_I / ((U0_I * U0_I))) * (dU[0][d]);
            wg[d] = (1.0 / U0_I) * (dU[3][d]) - (U3_I / ((U0_I * U0_I))) * (dU[0][d]);
        }
        for (int d = 0; d < 3; ++d)
            tg[d] = (1.0 / U0_I) * (dU[4][d]) - (U4_I / ((U0_I * U0_I))) * (dU[0][d]) - (U1_I / U0_I) * ug[d]
                    - (U2_I / U0_I) * vg[d] - (U3_I / U0_I) * wg[d];
        const double u_zeta = ug[0], u_eta = ug[1], u_xi = ug[2];
        const double v_zeta = vg[0], v_eta = vg[1], v_xi = vg[2];
        const double w_zeta = wg[0], w_eta = wg[1], w_xi = wg[2];
        const double t_zeta = tg[0], t_eta = tg[1], t_xi = tg[2];

        /* Viscous.cl:455-458 — temperature and Sutherland viscosity */
        const double d_temp = (1.4 * Mach * Mach * 0.4 * ((U4_I / U0_I) - 0.5 * (((U1_I * U1_I) / (U0_I * U0_I)) + ((U2_I * U2_I) / (U0_I * U0_I)) + ((U3_I * U3_I) / (U0_I * U0_I)))));
        const double d_temp1 = sqrt(d_temp) * sqrt(d_temp) * sqrt(d_temp);
        const double mu = (d_temp1) * ((1.0 + (110.4 / Free_t)) / (d_temp + (110.4 / Free_t)));

        /* Viscous.cl:461-463 — heat flux */
        const double qz = (-1.0) * (mu * 1.4 / (Pr * Reyl)) * (t_zeta * zeta_x + t_eta * eta_x + t_xi * xi_x);
        const double qe = (-1.0) * (mu * 1.4 / (Pr * Reyl)) * (t_zeta * zeta_y + t_eta * eta_y + t_xi * xi_y);
        const double qx = (-1.0) * (mu * 1.4 / (Pr * Reyl)) * (t_zeta * zeta_z + t_eta * eta_z + t_xi * xi_z);

        /* Viscous.cl:466-476 — Stokes stresses */
        const double tauzz = (mu / Reyl) * ((4.0 / 3.0) * (u_zeta * zeta_x + u_eta * eta_x + u_xi * xi_x) - (2.0 / 3.0) * (v_zeta * zeta_y + v_eta * eta_y + v_xi * xi_y) - (2.0 / 3.0) * (zeta_z * w_zeta + eta_z * w_eta + xi_z * w_xi));
        const double tauee = (mu / Reyl) * ((4.0 / 3.0) * (v_zeta * zeta_y + v_eta * eta_y + v_xi * xi_y) - (2.0 / 3.0) * (u_zeta * zeta_x + u_eta * eta_x + u_xi * xi_x) - (2.0 / 3.0) * (zeta_z * w_zeta + eta_z * w_eta + xi_z * w_xi));
        const double tauxx = (mu / Reyl) * ((4.0 / 3.0) * (w_zeta * zeta_z + w_eta * eta_z + w_xi * xi_z) - (2.0 / 3.0) * (u_zeta * zeta_x + u_eta * eta_x + u_xi * xi_x) - (2.0 / 3.0) * (zeta_y * v_zeta + eta_y * v_eta + xi_y * v_xi));
        const double tauze = (mu / Reyl) * (u_zeta * zeta_y + u_eta * eta_y + u_xi * xi_y + v_zeta * zeta_x + v_eta * eta_x + v_xi * xi_x);
        const double tauzx = (mu / Reyl) * (zeta_z * u_zeta + eta_z * u_eta + xi_z * u_xi + zeta_x * w_zeta + eta_x * w_eta + xi_x * w_xi);
        const double tauex = (mu / Reyl) * (zeta_z * v_zeta + eta_z * v_eta + xi_z * v_xi + zeta_y * w_zeta + eta_y * w_eta + xi_y * w_xi);

        /* Viscous.cl:490-558 — mixed-time-scale SGS model */
        const double C_MTS = prm->c_mts, C_T = prm->c_t;
        const double a_x = ((U1_I / U0_I) - ((0.25 * U_C[idx[0][2]].N1 + 0.5 * U1_I + 0.25 * U_C[idx[0][4]].N1) / (0.25 * U_C[idx[0][2]].N0 + 0.5 * U0_I + 0.25 * U_C[idx[0][4]].N0)));
        const double a_y = ((U2_I / U0_I) - ((0.25 * U_C[idx[1][2]].N2 + 0.5 * U2_I + 0.25 * U_C[idx[1][4]].N2) / (0.25 * U_C[idx[1][2]].N0 + 0.5 * U0_I + 0.25 * U_C[idx[1][4]].N0)));
        const double a_z = ((U3_I / U0_I) - ((0.25 * U_C[idx[2][2]].N3 + 0.5 * U3_I + 0.25 * U_C[idx[2][4]].N3) / (0.25 * U_C[idx[2][2]].N0 + 0.5 * U0_I + 0.25 * U_C[idx[2][4]].N0)));
        const double KS_x = pow(a_x, 2.0), KS_y = pow(a_y, 2.0), KS_z = pow(a_z, 2.0);
        const double stress_xy = (0.5 * (u_zeta * zeta_y + u_eta * eta_y + u_xi * xi_y + v_zeta * zeta_x + v_eta * eta_x + v_xi * xi_x));
        const double stress_yz = (0.5 * (v_zeta * zeta_z + v_eta * eta_z + v_xi * xi_z + w_zeta * zeta_y + w_eta * eta_y + w_xi * xi_y));
        const double stress_xz = (0.5 * (w_zeta * zeta_x + w_eta * eta_x + w_xi * xi_x + u_zeta * zeta_z + u_eta * eta_z + u_xi * xi_z));
        const double stress_xx = (u_zeta * zeta_x + u_eta * eta_x + u_xi * xi_x);
        const double stress_yy = (v_zeta * zeta_y + v_eta * eta_y + v_xi * xi_y);
        const double stress_zz = (w_zeta * zeta_z + w_eta * eta_z + w_xi * xi_z);
        const double mod_stress = sqrt(2.0 * ((stress_xy * stress_xy + stress_yz * stress_yz + stress_xz * stress_xz) + (2.0 / 3.0) * (stress_xx * stress_xx + stress_yy * stress_yy + stress_zz * stress_zz)));
        const double DEL = (1.0 / cbrt(DET_I));
        double EDDY_x, EDDY_y, EDDY_z;
        if (prm->sgs == 1) {
            /* Smagorinsky option: nu_t = (c_s det^(1/3))^2 |S|, |S| = sqrt(2 S_ij S_ij), carried as
               nu_t / det because the SGS stresses below take U0_I = det * rho */
            const double s2 = 2.0 * (stress_xx * stress_xx + stress_yy * stress_yy + stress_zz * stress_zz) +
                              4.0 * (stress_xy * stress_xy + stress_yz * stress_yz + stress_xz * stress_xz);
            EDDY_x = EDDY_y = EDDY_z = (prm->c_s * prm->c_s) * DEL * sqrt(s2);
        } else if (mod_stress != 0.0) {
            const double Rx = (DEL * (mod_stress)) / sqrt(KS_x);
            const double Ry = (DEL * (mod_stress)) / sqrt(KS_y);
            const double Rz = (DEL * (mod_stress)) / sqrt(KS_z);
            EDDY_x = (C_MTS / (1.0 + (1.0 * Rx / (C_T)))) * DEL * sqrt(KS_x);
            EDDY_y = (C_MTS / (1.0 + (1.0 * Ry / (C_T)))) * DEL * sqrt(KS_y);
            EDDY_z = (C_MTS / (1.0 + (1.0 * Rz / (C_T)))) * DEL * sqrt(KS_z);
        } else {
            EDDY_x = (C_MTS) * DEL * sqrt(KS_x);
            EDDY_y = (C_MTS) * DEL * sqrt(KS_y);
            EDDY_z = (C_MTS) * DEL * sqrt(KS_z);
        }
        /* the SGS stresses use U0_I = det*rho, as in Viscous.cl:529-535 */
        const double TAU_SGS_XY = -2.0 * U0_I * EDDY_x * (stress_xy);
        const double TAU_SGS_YZ = -2.0 * U0_I * EDDY_y * (stress_yz);
        const double TAU_SGS_XZ = -2.0 * U0_I * EDDY_z * (stress_xz);
        const double TAU_SGS_XX = -2.0 * U0_I * EDDY_x * stress_xx;
        const double TAU_SGS_YY = -2.0 * U0_I * EDDY_y * stress_yy;
        const double TAU_SGS_ZZ = -2.0 * U0_I * EDDY_z * stress_zz;
        const double Q_x = (1.0 / (1.4 * 0.4 * Mach * Mach)) * (-1.0) * (EDDY_x) * (t_zeta * zeta_x + t_eta * eta_x + t_xi * xi_x);
        const double Q_y = (1.0 / (1.4 * 0.4 * Mach * Mach)) * (-1.0) * (EDDY_y) * (t_zeta * zeta_y + t_eta * eta_y + t_xi * xi_y);
        const double Q_z = (1.0 / (1.4 * 0.4 * Mach * Mach)) * (-1.0) * (EDDY_z) * (t_zeta * zeta_z + t_eta * eta_z + t_xi * xi_z);
        const double DEL_X = TAU_SGS_XX * fl.u + TAU_SGS_XY * fl.u + TAU_SGS_XZ * fl.u;
        const double DEL_Y = TAU_SGS_XY * fl.v + TAU_SGS_YY * fl.v + TAU_SGS_YZ * fl.v;
        const double DEL_Z = TAU_SGS_XZ * fl.w + TAU_SGS_YZ * fl.w + TAU_SGS_ZZ * fl.w;
        double H_SGS_X = Q_x + DEL_X, H_SGS_Y = Q_y + DEL_Y, H_SGS_Z = Q_z + DEL_Z;
        double D_SGS_X = 0.4 * Q_x, D_SGS_Y = 0.4 * Q_y, D_SGS_Z = 0.4 * Q_z;
        if (prm->sgs == 1) {
            /* Smagorinsky option: heat flux -1.4 rho nu_t / Pr grad e (rho nu_t = U0_I * EDDY) and the
               work of the SGS stresses */
            const double kS = -(1.4 / prm->pr) * (U0_I * EDDY_x);
            H_SGS_X = kS * (t_zeta * zeta_x + t_eta * eta_x + t_xi * xi_x) +
                      (fl.u * TAU_SGS_XX + fl.v * TAU_SGS_XY + fl.w * TAU_SGS_XZ);
            H_SGS_Y = kS * (t_zeta * zeta_y + t_eta * eta_y + t_xi * xi_y) +
                      (fl.u * TAU_SGS_XY + fl.v * TAU_SGS_YY + fl.w * TAU_SGS_YZ);
            H_SGS_Z = kS * (t_zeta * zeta_z + t_eta * eta_z + t_xi * xi_z) +
                      (fl.u * TAU_SGS_XZ + fl.v * TAU_SGS_YZ + fl.w * TAU_SGS_ZZ);
            D_SGS_X = D_SGS_Y = D_SGS_Z = 0.0;
        }

        /* Viscous.cl:560-595 — Cartesian inviscid and viscous fluxes */
        const double rho = fl.rho, u = fl.u, v = fl.v, w = fl.w, p = fl.p;
        const double Et = ((p / 0.4) + 0.5 * rho * (u * u + v * v + w * w));
        const double tF[5] = {rho * u, rho * u * u + p, rho * v * u, rho * u * w, ((Et) + p) * u};
        const double tE[5] = {rho * v, rho * v * u, rho * v * v + p, rho * v * w, ((Et) + p) * v};
        const double tG[5] = {rho * w, rho * u * w, rho * v * w, rho * w * w + p, ((Et) + p) * w};
        const double tFv[5] = {0.0, (-1.0) * tauzz + TAU_SGS_XX, (-1.0) * tauze + TAU_SGS_XY, (-1.0) * tauzx + TAU_SGS_XZ,
                               (((-1.0) * u * tauzz - v * tauze - w * tauzx) + qz + H_SGS_X + D_SGS_X)};
        const double tEv[5] = {0.0, (-1.0) * tauze + TAU_SGS_XY, (-1.0) * tauee + TAU_SGS_YY, (-1.0) * tauex + TAU_SGS_YZ,
                               (((-1.0) * u * tauze - v * tauee - w * tauex) + qe + H_SGS_Y + D_SGS_Y)};
        const double tGv[5] = {0.0, (-1.0) * tauzx + TAU_SGS_XZ, (-1.0) * tauex + TAU_SGS_YZ, (-1.0) * tauxx + TAU_SGS_ZZ,
                               (((-1.0) * u * tauzx - v * tauex - w * tauxx) + qx + H_SGS_Z + D_SGS_Z)};

        /* Viscous.cl:597-631 — contravariant, det-scaled */
        double o[6][5];
        for (int c = 0; c < 5; ++c) {
            o[0][c] = DET_I * (zeta_x * tF[c] + zeta_y * tE[c] + zeta_z * tG[c]);
            o[1][c] = DET_I * (eta_x * tF[c] + eta_y * tE[c] + eta_z * tG[c]);
            o[2][c] = DET_I * (xi_x * tF[c] + xi_y * tE[c] + xi_z * tG[c]);
            o[3][c] = DET_I * (zeta_x * tFv[c] + zeta_y * tEv[c] + zeta_z * tGv[c]);
            o[4][c] = DET_I * (eta_x * tFv[c] + eta_y * tEv[c] + eta_z * tGv[c]);
            o[5][c] = DET_I * (xi_x * tFv[c] + xi_y * tEv[c] + xi_z * tGv[c]);
        }
        orc_vec5 *outs[6] = {&F_C[g], &E_C[g], &G_C[g], &Fv_C[g], &Ev_C[g], &Gv_C[g]};
        for (int k = 0; k < 6; ++k) {
            outs[k]->N0 = o[k][0]; outs[k]->N1 = o[k][1]; outs[k]->N2 = o[k][2];
            outs[k]->N3 = o[k][3]; outs[k]->N4 = o[k][4];
        }
        FLOW[g].mu = mu;                                                   /* Viscous.cl:641 */
    }
}

/* ------------------------------------------------------------------------- */
/* Weno.cl — one face of one node                                            */
/* ------------------------------------------------------------------------- */

typedef struct {
    double qpos[5], qneg[5];    /* det_f-scaled interface states (Weno.cl:2657-2727) */
    double cpos[5], cneg[5];    /* det_f * (k . (F, E, G)) of each state (Weno.cl:2585-2655) */
    double lpos, lneg;          /* lambda_5 of each state (Weno.cl:2511-2581) */
} orc_face;

/*
 * Face between stencil nodes s[2] (left) and s[3] (right); s[0..5] = offsets -2..+3
 * from the left node. roe = (u, v, w, p, h, rho) of that face, k = the face's
 * (unnormalised) metric row, detf = half-point det.
 * Weno.cl:336-455 (eigenvectors and projection), :1056-1279 (WCNS, repeated for
 * the five fields at :1280-2187), :2188-2287 (back projection), :2291-2507
 * (Euler fluxes), :2511-2581 (eigenvalues), :2585-2727 (contravariant and det).
 */
static void face_eigen(const double roe[6], const double k[3], double R[5][5], double L[5][5])
{
    const double ru = roe[0], rv = roe[1], rw = roe[2], rh = roe[4], rr = roe[5];
    const double kk = sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
    const double roe_a = sqrt(0.4 * (rh - 0.5 * (ru * ru + rv * rv + rw * rw)));
    const double kx = k[0] / kk, ky = k[1] / kk, kz = k[2] / kk;
    const double theta = kx * ru + ky * rv + kz * rw;
    const double phi_sq = 0.5 * 0.4 * (ru * ru + rv * rv + rw * rw);
    const double alpha = rr / (sqrt(2.0) * roe_a);
    const double beta = 1.0 / (sqrt(2.0) * rr * roe_a);
    const double aa = roe_a * roe_a;

    R[0][0] = kx; R[0][1] = ky; R[0][2] = kz; R[0][3] = alpha; R[0][4] = alpha;
    R[1][0] = kx * ru; R[1][1] = ky * ru - kz * rr; R[1][2] = kz * ru + ky * rr;
    R[1][3] = alpha * (ru + kx * roe_a); R[1][4] = alpha * (ru - kx * roe_a);
    R[2][0] = kx * rv + kz * rr; R[2][1] = ky * rv; R[2][2] = kz * rv - kx * rr;
    R[2][3] = alpha * (rv + ky * roe_a); R[2][4] = alpha * (rv - ky * roe_a);
    R[3][0] = kx * rw - ky * rr; R[3][1] = ky * rw + kx * rr; R[3][2] = kz * rw;
    R[3][3] = alpha * (rw + kz * roe_a); R[3][4] = alpha * (rw - kz * roe_a);
    R[4][0] = ((kx * phi_sq) / 0.4) + rr * (kz * rv - ky * rw);
    R[4][1] = ((ky * phi_sq) / 0.4) + rr * (kx * rw - kz * ru);
    R[4][2] = ((kz * phi_sq) / 0.4) + rr * (ky * ru - kx * rv);
    R[4][3] = alpha * (((phi_sq + roe_a * roe_a) / (0.4)) + theta * roe_a);
    R[4][4] = alpha * (((phi_sq + roe_a * roe_a) / (0.4)) - theta * roe_a);

    L[0][0] = kx * (1.0 - (phi_sq / (aa))) - ((kz * rv - ky * rw) / rr);
    L[0][1] = kx * 0.4 * ru / (aa);
    L[0][2] = (kx * 0.4 * rv / (aa)) + (kz / rr);
    L[0][3] = (kx * 0.4 * rw / (aa)) - (ky / rr);
    L[0][4] = -kx * 0.4 / (aa);
    L[1][0] = ky * (1.0 - (phi_sq / (aa))) - ((kx * rw - kz * ru) / rr);
    L[1][1] = (ky * 0.4 * ru / (aa)) - kz / rr;
    L[1][2] = ky * 0.4 * rv / (aa);
    L[1][3] = (ky * 0.4 * rw / (aa)) + (kx / rr);
    L[1][4] = -ky * 0.4 / (aa);
    L[2][0] = kz * (1.0 - (phi_sq / (aa))) - ((ky * ru - kx * rv) / rr);
    L[2][1] = kz * 0.4 * ru / (aa) + (ky / rr);
    L[2][2] = kz * 0.4 * rv / (aa) - (kx / rr);
    L[2][3] = kz * 0.4 * rw / (aa);
    L[2][4] = -kz * 0.4 / (aa);
    L[3][0] = beta * (phi_sq - theta * roe_a);
    L[3][1] = -beta * (0.4 * ru - kx * roe_a);
    L[3][2] = -beta * (0.4 * rv - ky * roe_a);
    L[3][3] = -beta * (0.4 * rw - kz * roe_a);
    L[3][4] = beta * 0.4;
    L[4][0] = beta * (phi_sq + theta * roe_a);
    L[4][1] = -beta * (0.4 * ru + kx * roe_a);
    L[4][2] = -beta * (0.4 * rv + ky * roe_a);
    L[4][3] = -beta * (0.4 * rw + kz * roe_a);
    L[4][4] = beta * 0.4;
}

/* WCNS interpolation of one characteristic field, both sides (Weno.cl:1056-1279) */
static void wcns_pair(const double Wc[6], double Epsilon, double *pos, double *neg)
{
    const double m2 = Wc[0], m1 = Wc[1], z0 = Wc[2], p1 = Wc[3], p2 = Wc[4], p3 = Wc[5];
    double hp[3], hm[3], ISp[3], ISm[3], wp[3], wm[3], Wp[3], Wm[3];
    hp[0] = (15.0 / 8.0) * p1 - (5.0 / 4.0) * p2 + (3.0 / 8.0) * p3;
    hp[1] = (3.0 / 8.0) * z0 + (3.0 / 4.0) * p1 - (1.0 / 8.0) * p2;
    hp[2] = (-1.0 / 8.0) * m1 + (3.0 / 4.0) * z0 + (3.0 / 8.0) * p1;
    hm[0] = (3.0 / 8.0) * z0 + (3.0 / 4.0) * p1 - (1.0 / 8.0) * p2;
    hm[1] = (-1.0 / 8.0) * m1 + (3.0 / 4.0) * z0 + (3.0 / 8.0) * p1;
    hm[2] = (3.0 / 8.0) * m2 - (5.0 / 4.0) * m1 + (15.0 / 8.0) * z0;
    ISm[2] = (13.0 / 12.0) * (pow(m2 - 2.0 * m1 + z0, 2)) + (1.0 / 4.0) * (pow(m2 - 4.0 * m1 + 3.0 * z0, 2));
    ISm[1] = (13.0 / 12.0) * (pow(m1 - 2.0 * z0 + p1, 2)) + (1.0 / 4.0) * (pow(m1 - p1, 2));
    ISm[0] = (13.0 / 12.0) * (pow(z0 - 2.0 * p1 + p2, 2)) + (1.0 / 4.0) * (pow(3.0 * z0 - 4.0 * p1 + p2, 2));
    ISp[2] = (13.0 / 12.0) * (pow(m1 - 2.0 * z0 + p1, 2)) + (1.0 / 4.0) * (pow(m1 - 4.0 * z0 + 3.0 * p1, 2));
    ISp[1] = (13.0 / 12.0) * (pow(z0 - 2.0 * p1 + p2, 2)) + (1.0 / 4.0) * (pow(z0 - p2, 2));
    ISp[0] = (13.0 / 12.0) * (pow(p1 - 2.0 * p2 + p3, 2)) + (1.0 / 4.0) * (pow(3.0 * p1 - 4.0 * p2 + p3, 2));
    wp[0] = (1.0 / 16.0) / (pow(Epsilon + (ISp[0]), 2.0));
    wp[1] = (5.0 / 8.0) / (pow(Epsilon + (ISp[1]), 2.0));
    wp[2] = (5.0 / 16.0) / (pow(Epsilon + (ISp[2]), 2.0));
    wm[0] = (5.0 / 16.0) / (pow(Epsilon + (ISm[0]), 2.0));
    wm[1] = (5.0 / 8.0) / (pow(Epsilon + (ISm[1]), 2.0));
    wm[2] = (1.0 / 16.0) / (pow(Epsilon + (ISm[2]), 2.0));
    for (int q = 0; q < 3; ++q) {
        Wp[q] = wp[q] / (wp[0] + wp[1] + wp[2]);
        Wm[q] = wm[q] / (wm[0] + wm[1] + wm[2]);
    }
    *pos = Wp[0] * hp[0] + Wp[1] * hp[1] + Wp[2] * hp[2];
    *neg = Wm[0] * hm[0] + Wm[1] * hm[1] + Wm[2] * hm[2];
}

void orc_wcns_pair(const double W[6], double eps, double *pos, double *neg) { wcns_pair(W, eps, pos, neg); }

void orc_face_eigen(const double roe[6], const double k[3], double L[25], double R[25])
{
    double l[5][5], r[5][5];
    face_eigen(roe, k, r, l);
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) { L[5 * i + j] = l[i][j]; R[5 * i + j] = r[i][j]; }
}

static void weno_face(const orc_vec5 *s[6], const double roe[6], const double k[3], double detf, double Epsilon,
                      orc_face *out)
{
    double R[5][5], L[5][5];
    face_eigen(roe, k, R, L);

    /* W[m][c] = characteristic field c at stencil point m (Weno.cl:415-455) */
    double W[6][5];
    for (int m = 0; m < 6; ++m)
        for (int c = 0; c < 5; ++c)
            W[m][c] = L[c][0] * s[m]->N0 + L[c][1] * s[m]->N1 + L[c][2] * s[m]->N2 + L[c][3] * s[m]->N3 + L[c][4] * s[m]->N4;

    double cpos[5], cneg[5];
    for (int c = 0; c < 5; ++c) {
        const double Wc[6] = {W[0][c], W[1][c], W[2][c], W[3][c], W[4][c], W[5][c]};
        wcns_pair(Wc, Epsilon, &cpos[c], &cneg[c]);
    }

    double qp[5], qn[5];
    for (int m = 0; m < 5; ++m) {
        qp[m] = R[m][0] * (cpos[0]) + R[m][1] * (cpos[1]) + R[m][2] * (cpos[2]) + R[m][3] * (cpos[3]) + R[m][4] * (cpos[4]);
        qn[m] = R[m][0] * (cneg[0]) + R[m][1] * (cneg[1]) + R[m][2] * (cneg[2]) + R[m][3] * (cneg[3]) + R[m][4] * (cneg[4]);
    }

    for (int side = 0; side < 2; ++side) {
        const double *Q = side == 0 ? qp : qn;
        const double ke = (pow(Q[1], 2) / Q[0]) + (pow(Q[2], 2) / Q[0]) + (pow(Q[3], 2) / Q[0]);
        double Fx[5], Fy[5], Fz[5];
        Fx[0] = Q[1];
        Fx[1] = (pow(Q[1], 2) / Q[0]) + 0.4 * (Q[4] - 0.5 * (ke));
        Fx[2] = (Q[1] * Q[2]) / Q[0];
        Fx[3] = (Q[1] * Q[3]) / Q[0];
        Fx[4] = (1.4 * Q[4] - 0.2 * (ke)) * (Q[1] / Q[0]);
        Fy[0] = Q[2];
        Fy[1] = (Q[1] * Q[2]) / Q[0];
        Fy[2] = (pow(Q[2], 2) / Q[0]) + 0.4 * (Q[4] - 0.5 * (ke));
        Fy[3] = (Q[2] * Q[3]) / Q[0];
        Fy[4] = (1.4 * Q[4] - 0.2 * (ke)) * (Q[2] / Q[0]);
        Fz[0] = Q[3];
        Fz[1] = (Q[1] * Q[3]) / Q[0];
        Fz[2] = (Q[2] * Q[3]) / Q[0];
        Fz[3] = (pow(Q[3], 2) / Q[0]) + 0.4 * (Q[4] - 0.5 * (ke));
        Fz[4] = (1.4 * Q[4] - 0.2 * (ke)) * (Q[3] / Q[0]);
        const double lam = (k[0] * (Q[1] / Q[0]) + k[1] * (Q[2] / Q[0]) + k[2] * (Q[3] / Q[0]))
                           + sqrt(1.4 * 0.4 * ((Q[4] / Q[0]) - ((pow(Q[1], 2) + pow(Q[2], 2) + pow(Q[3], 2)) / (2.0 * pow(Q[0], 2))))
                                  * (k[0] * k[0] + k[1] * k[1] + k[2] * k[2]));
        double *cc = side == 0 ? out->cpos : out->cneg;
        double *qq = side == 0 ? out->qpos : out->qneg;
        for (int c = 0; c < 5; ++c) {
            cc[c] = detf * (k[0] * Fx[c] + k[1] * Fy[c] + k[2] * Fz[c]);
            qq[c] = detf * Q[c];
        }
        if (side == 0) out->lpos = lam; else out->lneg = lam;
    }
}

/* LLF (Weno.cl:2859-2947) plus the explicit midpoint correction (Weno.cl:2951-3058).
 * nf[0..5] = node fluxes at offsets -2..+3 from the face's left node. */
static void face_flux(const orc_face *f, const orc_vec5 *nf[6], double out[5])
{
    double alpha;
    if (fabs(f->lpos) >= fabs(f->lneg)) alpha = fabs(f->lpos);
    else alpha = fabs(f->lneg);
    for (int c = 0; c < 5; ++c) {
        const double llf = 0.5 * (f->cpos[c] + f->cneg[c] - alpha * (f->qpos[c] - f->qneg[c]));
        const double a = v5(nf[0], c), b = v5(nf[1], c), z = v5(nf[2], c), d = v5(nf[3], c), e = v5(nf[4], c), g = v5(nf[5], c);
        const double d2 = (1.0 / 48.0) * (-5.0 * a + 39.0 * b - 34.0 * z - 34.0 * d + 39.0 * e - 5.0 * g);
        const double d4 = (1.0 / 2.0) * (a - 3.0 * b + 2.0 * z + 2.0 * d - 3.0 * e + g);
        out[c] = llf - (1.0 / 24.0) * d2 + (7.0 / 5760.0) * d4;
    }
}

/* ------------------------------------------------------------------------- */
/* Weno.cl:36-3137                                                           */
/* ------------------------------------------------------------------------- */

void orc_weno(int64_t n_launch, int j, double Mach, double del_t, const orc_deter *deter, const double *det,
              orc_vec5 *U0, orc_vec5 *U1, orc_vec5 *U2, const orc_grid *Grid_P, const orc_vec5 *Q_C,
              const orc_metric *metric, const orc_roe *ROE_AVER, const orc_vec5 *Fv, const orc_vec5 *Ev,
              const orc_vec5 *Gv, const orc_vec5 *F, const orc_vec5 *E, const orc_vec5 *G, orc_flow *FLOW,
              const orc_params *prm)
{
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < n_launch; ++g) {
        int64_t idx[3][7];
        int loc[3][7];
        for (int d = 0; d < 3; ++d) {
            walk(Grid_P, g, d, idx[d]);
            for (int s = 0; s < 7; ++s) loc[d][s] = Grid_P[idx[d][s]].loc;
        }
        const orc_vec5 *nodeF[3] = {F, E, G};
        const orc_vec5 *nodeV[3] = {Fv, Ev, Gv};
        /* deter order ip, im, jp, jm, kp, km; ROE order ip, jp, kp, im, jm, km; metric faces ip, im, jp, jm, kp, km */
        const double detp[3] = {deter[g].ip, deter[g].jp, deter[g].kp};
        const double detm[3] = {deter[g].im, deter[g].jm, deter[g].km};
        double dflux[3][5], dvisc[3][5];
        for (int d = 0; d < 3; ++d) {
            const double *kp = &metric[g].face[2 * d][3 * d];      /* zeta_?ip, eta_?jp, xi_?kp */
            const double *km = &metric[g].face[2 * d + 1][3 * d];  /* zeta_?im, eta_?jm, xi_?km */
            const orc_vec5 *sp[6], *sm[6], *fp[6], *fm[6];
            for (int s = 0; s < 6; ++s) {
                sp[s] = &Q_C[idx[d][s + 1]];        /* plus face: offsets -2..+3 from g */
                sm[s] = &Q_C[idx[d][s]];            /* minus face: offsets -3..+2 from g */
                fp[s] = &nodeF[d][idx[d][s + 1]];
                fm[s] = &nodeF[d][idx[d][s]];
            }
            orc_face fpl, fmi;
            weno_face(sp, ROE_AVER[g].f[d], kp, detp[d], prm->eps, &fpl);
            weno_face(sm, ROE_AVER[g].f[3 + d], km, detm[d], prm->eps, &fmi);
            double hp[5], hm[5];
            face_flux(&fpl, fp, hp);
            face_flux(&fmi, fm, hm);
            for (int c = 0; c < 5; ++c) {
                dflux[d][c] = hp[c] - hm[c];                                   /* Weno.cl:3060-3076 */
                double f[7];
                for (int s = 0; s < 7; ++s) f[s] = v5(&nodeV[d][idx[d][s]], c);
                dvisc[d][c] = loc_deriv(f, loc[d], 1);                         /* Weno.cl:2732-2856 */
            }
        }
        double L[5];
        for (int c = 0; c < 5; ++c)
            L[c] = (-1.0) * (dflux[0][c] + dflux[1][c] + dflux[2][c] + dvisc[0][c] + dvisc[1][c] + dvisc[2][c]);

        /* Weno.cl:3084-3126 — SSP-RK3 stage j */
        double fu[5];
        double *u0 = &U0[g].N0, *u1 = &U1[g].N0, *u2 = &U2[g].N0;
        if (j == 0) {
            for (int c = 0; c < 5; ++c) u1[c] = u0[c] + del_t * L[c];
            for (int c = 0; c < 5; ++c) fu[c] = (1.0 / det[g]) * u1[c];
        } else if (j == 1) {
            for (int c = 0; c < 5; ++c) u2[c] = (3.0 / 4.0) * u0[c] + (1.0 / 4.0) * u1[c] + 0.25 * del_t * L[c];
            for (int c = 0; c < 5; ++c) fu[c] = (1.0 / det[g]) * u2[c];
        } else {
            for (int c = 0; c < 5; ++c) u0[c] = (u0[c] / 3.0) + ((2.0 / 3.0) * u2[c]) + (2.0 / 3.0) * (del_t * L[c]);
            for (int c = 0; c < 5; ++c) fu[c] = (1.0 / det[g]) * u0[c];
        }
        /* Weno.cl:3128-3135 — primitives */
        orc_flow *fl = &FLOW[g];
        fl->rho = fu[0];
        fl->u = fu[1] / fu[0];
        fl->v = fu[2] / fu[0];
        fl->w = fu[3] / fu[0];
        fl->e = (fu[4] / fu[0]) - 0.5 * (pow(fl->u, 2.0) + pow(fl->v, 2.0) + pow(fl->w, 2.0));
        fl->p = fl->e * fl->rho * 0.4;
        fl->t = (1.4 * Mach * Mach) * (fl->p / fl->rho);
        fl->a = sqrt(1.4 * fl->p / fl->rho);
    }
}

