/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement (plain C99, double precision, no FMA contraction) of the
 * reference solver's per-RK-stage hot path: the three OpenCL kernels U_Q.cl,
 * Viscous.cl and Weno.cl, and the host stage sequence of Solver.cpp:356-541
 * (USE_GPU==1 data contract). Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker; the
 * product (libcfdx) never links or calls it.
 *
 * The same entry points are exported by oracle/_ref/libref.so, which is the
 * reference's own kernel source (U_Q.cl, Viscous.cl, Weno.cl) compiled as C
 * by oracle/Makefile where the reference tree is present; tests run both on
 * identical inputs.
 *
 * Struct layouts follow functions.h:49-112 byte for byte (packed).
 */
#ifndef CFDX_ORACLE_H
#define CFDX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct __attribute__((packed)) { double u, v, w, p, t, rho, mu, a, e; } orc_flow;      /* 72 B  functions.h:49-51 */
typedef struct __attribute__((packed)) { double N0, N1, N2, N3, N4; } orc_vec5;               /* 40 B  functions.h:53-55 */
typedef struct __attribute__((packed)) { double x, y, z; int N0, N1, N2, N3, N4, N5, ID, loc; } orc_grid; /* 56 B functions.h:57-61 */
/* ROE_AVERAGING functions.h:63-70: faces in the order ip, jp, kp, im, jm, km; per face (u, v, w, p, h, rho). */
typedef struct __attribute__((packed)) { double f[6][6]; } orc_roe;                          /* 288 B */
/* TRANSFORMED functions.h:93-98: 9 node metrics (zeta_x..xi_z) then for each face ip, im, jp, jm, kp, km the 9 metrics. */
typedef struct __attribute__((packed)) { double node[9]; double face[6][9]; } orc_metric;      /* 504 B */
typedef struct __attribute__((packed)) { double ip, im, jp, jm, kp, km; } orc_deter;           /* 48 B  functions.h:110-112 */

/* Constants the reference hard-codes in Viscous.cl:30-33, :490-491 and Weno.cl:59.
 * The reference values are Mach 1.5, Re 500, Pr 0.72, T_inf 273, C_MTS 0.03, C_T 10, eps 10e-8. */
typedef struct {
    double mach;      /* Viscous.cl:30 (temperature and SGS heat flux) */
    double reyl;      /* Viscous.cl:31 (INFINITY gives an inviscid stress) */
    double pr;        /* Viscous.cl:32 */
    double free_t;    /* Viscous.cl:33 (Sutherland reference temperature) */
    double c_mts;     /* Viscous.cl:490 */
    double c_t;       /* Viscous.cl:491 */
    double eps;       /* Weno.cl:59 */
    int32_t sgs;      /* 0: the reference's mixed-time-scale model; 1: Smagorinsky, nu_t = (c_s det^(1/3))^2 |S|
                         (no reference counterpart; the reference kernels compiled as C ignore it) */
    int32_t pad_;
    double c_s;
} orc_params;

/* U_Q.cl:36-125 for Gid = 0 .. n_launch-1 */
void orc_uq(int64_t n_launch, orc_vec5 *U_C, orc_vec5 *Q_C, const orc_grid *Grid_P,
            const double *det, orc_flow *FLOW, orc_roe *ROE_AVER);

/* Viscous.cl:23-643 */
void orc_viscous(int64_t n_launch, const orc_vec5 *U_C, orc_vec5 *F_C, orc_vec5 *E_C, orc_vec5 *G_C,
                 orc_vec5 *Fv_C, orc_vec5 *Ev_C, orc_vec5 *Gv_C, const orc_grid *Grid_P,
                 const orc_metric *metric, const double *det, orc_flow *FLOW, const orc_params *prm);

/* Weno.cl:36-3137 */
void orc_weno(int64_t n_launch, int j, double Mach, double del_t, const orc_deter *deter, const double *det,
              orc_vec5 *U0, orc_vec5 *U1, orc_vec5 *U2, const orc_grid *Grid_P, const orc_vec5 *Q_C,
              const orc_metric *metric, const orc_roe *ROE_AVER, const orc_vec5 *Fv, const orc_vec5 *Ev,
              const orc_vec5 *Gv, const orc_vec5 *F, const orc_vec5 *E, const orc_vec5 *G, orc_flow *FLOW,
              const orc_params *prm);

/* Solver.cpp:356-541 stage sequence (USE_GPU==1 contract), nsteps full RK3 steps, no halo exchange or
 * boundary fill (single periodic subdomain). Scratch arrays are caller-owned, each n_g entries. */
typedef struct {
    orc_vec5 *U0, *U1, *U2, *Q_C, *F, *E, *G, *Fv, *Ev, *Gv;
    orc_roe *ROE;
} orc_work;
void orc_solver_steps(int64_t n_g, int64_t n_sd, int nsteps, double Mach, double del_t,
                      const orc_grid *Grid_P, const double *det, const orc_metric *metric,
                      const orc_deter *deter, orc_flow *FLOW, orc_work *w, const orc_params *prm);

/* Unit-test hooks into the restatement's building blocks (liboracle.so only):
 * the WCNS pair of one characteristic field (W[0..5] at offsets -2..+3 from the
 * face's left node; Weno.cl:1056-1279), and the L / R eigenvector matrices of a
 * face (row-major 5x5; Weno.cl:336-412). */
void orc_wcns_pair(const double W[6], double eps, double *pos, double *neg);
void orc_face_eigen(const double roe[6], const double k[3], double L[25], double R[25]);

/* Identifies which implementation is loaded: 0 = restatement, 1 = reference kernels. */
int orc_kind(void);
/* Set the OpenMP thread count used by the loops above (0 = library default). */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
