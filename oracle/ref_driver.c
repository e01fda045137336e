/*
 * ORACLE — test infrastructure only (see oracle.h).
 * Drives the reference's own kernels (/root/reference/{U_Q,Viscous,Weno}.cl,
 * compiled as C99 by oracle/Makefile) over Gid = 0 .. n_launch-1, the role the
 * OpenCL runtime plays in Solver.cpp:377,440,521. The work-item index is a
 * thread-local variable read through the get_global_id(0) macro of the build
 * recipe. The constants of Viscous.cl:30-33,490-491 and Weno.cl:59 are fixed in
 * that source, so the params argument is ignored here: tests call this library
 * with the reference values only.
 */
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

_Thread_local unsigned oracle_gid;

/* kernel entry points as compiled from the .cl sources (struct layouts are the packed ones of oracle.h) */
void U_Q(orc_vec5 *U_C, orc_vec5 *Q_C, const orc_grid *Grid_P, const double *det, orc_flow *FLOW, orc_roe *ROE_AVER);
void Viscous(const orc_vec5 *U_C, orc_vec5 *F_C, orc_vec5 *E_C, orc_vec5 *G_C, orc_vec5 *Fv_C, orc_vec5 *Ev_C,
             orc_vec5 *Gv_C, const orc_grid *Grid_P, const orc_metric *metric, const double *det, orc_flow *FLOW);
void Weno(const int j, const double Mach, const double del_t, const orc_deter *deter, const double *det,
          orc_vec5 *U0, orc_vec5 *U1, orc_vec5 *U2, const orc_grid *Grid_P, const orc_vec5 *Q_C,
          const orc_metric *metric, const orc_roe *ROE_AVER, const orc_vec5 *Fv, const orc_vec5 *Ev,
          const orc_vec5 *Gv, const orc_vec5 *F, const orc_vec5 *E, const orc_vec5 *G, orc_flow *FLOW);

int orc_kind(void) { return 1; }

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void orc_uq(int64_t n_launch, orc_vec5 *U_C, orc_vec5 *Q_C, const orc_grid *Grid_P,
            const double *det, orc_flow *FLOW, orc_roe *ROE_AVER)
{
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < n_launch; ++g) {
        oracle_gid = (unsigned)g;
        U_Q(U_C, Q_C, Grid_P, det, FLOW, ROE_AVER);
    }
}

void orc_viscous(int64_t n_launch, const orc_vec5 *U_C, orc_vec5 *F_C, orc_vec5 *E_C, orc_vec5 *G_C,
                 orc_vec5 *Fv_C, orc_vec5 *Ev_C, orc_vec5 *Gv_C, const orc_grid *Grid_P,
                 const orc_metric *metric, const double *det, orc_flow *FLOW, const orc_params *prm)
{
    (void)prm;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < n_launch; ++g) {
        oracle_gid = (unsigned)g;
        Viscous(U_C, F_C, E_C, G_C, Fv_C, Ev_C, Gv_C, Grid_P, metric, det, FLOW);
    }
}

void orc_weno(int64_t n_launch, int j, double Mach, double del_t, const orc_deter *deter, const double *det,
              orc_vec5 *U0, orc_vec5 *U1, orc_vec5 *U2, const orc_grid *Grid_P, const orc_vec5 *Q_C,
              const orc_metric *metric, const orc_roe *ROE_AVER, const orc_vec5 *Fv, const orc_vec5 *Ev,
              const orc_vec5 *Gv, const orc_vec5 *F, const orc_vec5 *E, const orc_vec5 *G, orc_flow *FLOW,
              const orc_params *prm)
{
    (void)prm;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < n_launch; ++g) {
        oracle_gid = (unsigned)g;
        Weno(j, Mach, del_t, deter, det, U0, U1, U2, Grid_P, Q_C, metric, ROE_AVER, Fv, Ev, Gv, F, E, G, FLOW);
    }
}
