/*
 * ORACLE — test infrastructure only (see oracle.h).
 * Solver.cpp:356-541 stage sequence under the USE_GPU==1 data contract, linked
 * into both liboracle.so (restatement) and _ref/libref.so (reference kernels):
 *   j = 0: U_Q writes U0 = det*Q(FLOW) (kept as the RK base, Solver.cpp:385-388);
 *   j > 0: U_Q's U output is discarded and the viscous kernel reads the raw stage
 *          vector U1 / U2 instead (Solver.cpp:427-435);
 *   Weno writes U1 (j=0), U2 (j=1), U0 (j=2) and FLOW (Weno.cl:3084-3135).
 * No halo exchange or boundary fill: one periodic subdomain (Solver.cpp:576-680 skipped).
 */
#include "oracle.h"

void orc_solver_steps(int64_t n_g, int64_t n_sd, int nsteps, double Mach, double del_t,
                      const orc_grid *Grid_P, const double *det, const orc_metric *metric,
                      const orc_deter *deter, orc_flow *FLOW, orc_work *w, const orc_params *prm)
{
    for (int it = 0; it < nsteps; ++it) {
        for (int j = 0; j < 3; ++j) {
            /* at j > 0 the U output lands in F, which the viscous kernel rewrites next */
            orc_uq(n_g, j == 0 ? w->U0 : w->F, w->Q_C, Grid_P, det, FLOW, w->ROE);
            const orc_vec5 *uv = j == 0 ? w->U0 : j == 1 ? w->U1 : w->U2;
            orc_viscous(n_g, uv, w->F, w->E, w->G, w->Fv, w->Ev, w->Gv, Grid_P, metric, det, FLOW, prm);
            orc_weno(n_sd, j, Mach, del_t, deter, det, w->U0, w->U1, w->U2, Grid_P, w->Q_C, metric, w->ROE,
                     w->Fv, w->Ev, w->Gv, w->F, w->E, w->G, FLOW, prm);
        }
    }
}
