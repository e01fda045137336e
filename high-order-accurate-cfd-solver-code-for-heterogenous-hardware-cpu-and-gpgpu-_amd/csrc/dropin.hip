// Drop-in kernels: the reference's U_Q, Viscous and Weno with the same arguments,
// packed AoS layouts and per-node semantics (one work-item per node, index 0 the
// dummy node), for meshes given by a GRID neighbour table (any numbering, ghosts
// and loc codes included). Packed records are read as plain double / int32 arrays
// at their byte offsets so every access is naturally aligned.
#include <hip/hip_runtime.h>

#include "ctx.hpp"
#include "physics.hpp"

using namespace cfdx::dev;

namespace {

// record strides and field offsets (in elements) of the packed layouts, functions.h:49-112
constexpr int FLOW_N = 9, F_U = 0, F_V = 1, F_W = 2, F_P = 3, F_T = 4, F_RHO = 5, F_MU = 6, F_A = 7, F_E = 8;
constexpr int GRID_N = 14, G_NB = 6, G_LOC = 13;     // int32 view: x,y,z fill slots 0..5
constexpr int ROE_N = 36, MET_N = 63, DET_N = 6;

__device__ __forceinline__ int nb_plus(const int32_t *G, int g, int d)
{
    // +i = N1, +j = N0, +k = N4
    return G[GRID_N * g + G_NB + (d == 0 ? 1 : d == 1 ? 0 : 4)];
}
__device__ __forceinline__ int nb_minus(const int32_t *G, int g, int d)
{
    // -i = N3, -j = N2, -k = N5
    return G[GRID_N * g + G_NB + (d == 0 ? 3 : d == 1 ? 2 : 5)];
}

// idx[0..6] = nodes at offsets -3..+3 along direction d (Viscous.cl:99-118)
__device__ __forceinline__ void walk(const int32_t *G, int g, int d, int (&idx)[7])
{
    idx[3] = g;
    int p = g, m = g;
#pragma unroll
    for (int s = 1; s <= 3; ++s) {
        p = nb_plus(G, p, d);
        m = nb_minus(G, m, d);
        idx[3 + s] = p;
        idx[3 - s] = m;
    }
}

// ---------------------------------------------------------------------------
// U_Q.cl:36-125
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_uq(double *__restrict__ U_C, double *__restrict__ Q_C, const int32_t *__restrict__ G,
                                            const double *__restrict__ det, const double *__restrict__ FLOW,
                                            double *__restrict__ ROE, int n)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const double *f = FLOW + (size_t)FLOW_N * g;
    const double rho = f[F_RHO], u = f[F_U], v = f[F_V], w = f[F_W], p = f[F_P];
    const double h = enthalpy(rho, u, v, w, p);
    // ROE slots ip, jp, kp, im, jm, km -> N1, N0, N4, N3, N2, N5 (U_Q.cl:51-56)
    const int slot_nb[6] = {1, 0, 4, 3, 2, 5};
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int nb = G[GRID_N * g + G_NB + slot_nb[s]];
        const double *fn = FLOW + (size_t)FLOW_N * nb;
        const double rn = fn[F_RHO], un = fn[F_U], vn = fn[F_V], wn = fn[F_W], pn = fn[F_P];
        double r[5];
        roe_average(rho, u, v, w, h, rn, un, vn, wn, enthalpy(rn, un, vn, wn, pn), r);
        double *ro = ROE + (size_t)ROE_N * g + 6 * s;
        ro[0] = r[0]; ro[1] = r[1]; ro[2] = r[2]; ro[4] = r[3]; ro[5] = r[4];   // roe_p is never written
    }
    double q[5];
    prim_to_q(rho, u, v, w, p, q);
    const double D = det[g];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        Q_C[(size_t)5 * g + c] = q[c];
        U_C[(size_t)5 * g + c] = D * q[c];
    }
}

// ---------------------------------------------------------------------------
// Viscous.cl:23-643
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(128) k_viscous(const double *__restrict__ U_C, double *__restrict__ Fo, double *__restrict__ Eo,
                                                 double *__restrict__ Go, double *__restrict__ Fvo, double *__restrict__ Evo,
                                                 double *__restrict__ Gvo, const int32_t *__restrict__ G,
                                                 const double *__restrict__ metric, const double *__restrict__ det,
                                                 double *__restrict__ FLOW, int n, ViscParams vp)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    double dU[5][3];
    double KS[3];
    double U[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) U[c] = U_C[(size_t)5 * g + c];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int idx[7];
        walk(G, g, d, idx);
        int loc[7];
#pragma unroll
        for (int s = 0; s < 7; ++s) loc[s] = G[GRID_N * idx[s] + G_LOC];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            double f[7];
#pragma unroll
            for (int s = 0; s < 7; ++s) f[s] = U_C[(size_t)5 * idx[s] + c];
            dU[c][d] = loc_derivative(f, loc[3], loc[1], loc[2], loc[4], loc[5], false);
        }
        // MTS scale-similarity term of momentum component d+1 along d (Viscous.cl:493-495)
        const double *um = U_C + (size_t)5 * idx[2], *up = U_C + (size_t)5 * idx[4];
        const double a = (U[1 + d] / U[0]) - ((0.25 * um[1 + d] + 0.5 * U[1 + d] + 0.25 * up[1 + d]) / (0.25 * um[0] + 0.5 * U[0] + 0.25 * up[0]));
        KS[d] = a * a;
    }
    double m[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) m[c] = metric[(size_t)MET_N * g + c];
    const double D = det[g];
    double *f = FLOW + (size_t)FLOW_N * g;
    const double rho = f[F_RHO], u = f[F_U], v = f[F_V], w = f[F_W], p = f[F_P];
    double fv[3][5];
    const double mu = viscous_node(dU, U, u, v, w, m, D, KS, vp, fv);
    double *outv[3] = {Fvo, Evo, Gvo};
    double *outi[3] = {Fo, Eo, Go};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double nf[5];
        node_flux(rho, u, v, w, p, m[3 * d], m[3 * d + 1], m[3 * d + 2], D, nf);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            outi[d][(size_t)5 * g + c] = nf[c];
            outv[d][(size_t)5 * g + c] = fv[d][c];
        }
    }
    f[F_MU] = mu;
}

// ---------------------------------------------------------------------------
// Weno.cl:36-3137
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(128) k_weno(int j, double Mach, double del_t, const double *__restrict__ deter,
                                              const double *__restrict__ det, double *__restrict__ U0, double *__restrict__ U1,
                                              double *__restrict__ U2, const int32_t *__restrict__ G,
                                              const double *__restrict__ Q_C, const double *__restrict__ metric,
                                              const double *__restrict__ ROE, const double *__restrict__ Fv,
                                              const double *__restrict__ Ev, const double *__restrict__ Gv,
                                              const double *__restrict__ F, const double *__restrict__ E,
                                              const double *__restrict__ Gf, double *__restrict__ FLOW, int n, double eps)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const double *nodeF[3] = {F, E, Gf};
    const double *nodeV[3] = {Fv, Ev, Gv};
    double L[5] = {0, 0, 0, 0, 0};
    double dvis[3][5];
#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
        int idx[7];
        walk(G, g, d, idx);
        double hp[5], hm[5];
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            // plus face: stencil offsets -2..+3 from g; minus face: -3..+2 (left node g-1)
            const int o = side == 0 ? 1 : 0;
            double q[6][5], nf[6][5];
#pragma unroll
            for (int s = 0; s < 6; ++s)
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    q[s][c] = Q_C[(size_t)5 * idx[s + o] + c];
                    nf[s][c] = nodeF[d][(size_t)5 * idx[s + o] + c];
                }
            // ROE slots ip, jp, kp, im, jm, km; metric faces ip, im, jp, jm, kp, km; deter ip, im, jp, jm, kp, km
            const double *ro = ROE + (size_t)ROE_N * g + 6 * (side == 0 ? d : 3 + d);
            const double roe[5] = {ro[0], ro[1], ro[2], ro[4], ro[5]};
            const double *k = metric + (size_t)MET_N * g + 9 + 9 * (2 * d + side) + 3 * d;
            const double detf = deter[(size_t)DET_N * g + 2 * d + side];
            if (side == 0) face_flux(q, roe, k[0], k[1], k[2], detf, eps, nf, hp);
            else face_flux(q, roe, k[0], k[1], k[2], detf, eps, nf, hm);
        }
        int loc[7];
#pragma unroll
        for (int s = 0; s < 7; ++s) loc[s] = G[GRID_N * idx[s] + G_LOC];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            double fv[7];
#pragma unroll
            for (int s = 0; s < 7; ++s) fv[s] = nodeV[d][(size_t)5 * idx[s] + c];
            dvis[d][c] = loc_derivative(fv, loc[3], loc[1], loc[2], loc[4], loc[5], true);
            L[c] += hp[c] - hm[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) L[c] = -(L[c] + dvis[0][c] + dvis[1][c] + dvis[2][c]);   // Weno.cl:3078-3082

    // Weno.cl:3084-3126 SSP-RK3 stage j
    double fu[5];
    const double idet = 1.0 / det[g];
    double *u0 = U0 + (size_t)5 * g, *u1 = U1 + (size_t)5 * g, *u2 = U2 + (size_t)5 * g;
    if (j == 0) {
#pragma unroll
        for (int c = 0; c < 5; ++c) { const double x = u0[c] + del_t * L[c]; u1[c] = x; fu[c] = idet * x; }
    } else if (j == 1) {
#pragma unroll
        for (int c = 0; c < 5; ++c) { const double x = 0.75 * u0[c] + 0.25 * u1[c] + 0.25 * del_t * L[c]; u2[c] = x; fu[c] = idet * x; }
    } else {
#pragma unroll
        for (int c = 0; c < 5; ++c) { const double x = (u0[c] / 3.0) + ((2.0 / 3.0) * u2[c]) + (2.0 / 3.0) * (del_t * L[c]); u0[c] = x; fu[c] = idet * x; }
    }
    // Weno.cl:3128-3135 primitives
    double *fl = FLOW + (size_t)FLOW_N * g;
    const double rho = fu[0], ir = 1.0 / rho;
    const double u = fu[1] * ir, v = fu[2] * ir, w = fu[3] * ir;
    const double e = fu[4] * ir - 0.5 * (u * u + v * v + w * w);
    const double p = e * rho * 0.4;
    fl[F_RHO] = rho; fl[F_U] = u; fl[F_V] = v; fl[F_W] = w; fl[F_E] = e; fl[F_P] = p;
    fl[F_T] = (1.4 * Mach * Mach) * (p * ir);
    fl[F_A] = sqrt(1.4 * p * ir);
}

cfdx::dev::ViscParams visc_params(const cfdx_params *prm)
{
    cfdx_params d;
    if (!prm) { cfdx_params_default(&d); prm = &d; }
    return cfdx::dev::make_visc_params(prm->mach, prm->reyl, prm->pr, prm->free_t, prm->c_mts, prm->c_t, prm->sgs, prm->c_s);
}

}  // namespace

extern "C" int cfdx_uq(cfdx_ctx *ctx, cfdx_vec5 *U_C, cfdx_vec5 *Q_C, const cfdx_grid *Grid_P, const double *det,
                       cfdx_flow *FLOW, cfdx_roe *ROE_AVER, int64_t n_launch)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (!U_C || !Q_C || !Grid_P || !det || !FLOW || !ROE_AVER) return cfdx::fail(CFDX_E_INVALID, "cfdx_uq: null buffer");
    if (n_launch < 0 || n_launch > INT32_MAX) return cfdx::fail(CFDX_E_INVALID, "cfdx_uq: n_launch %lld out of range", (long long)n_launch);
    if (n_launch == 0) return CFDX_OK;
    hipLaunchKernelGGL(k_uq, dim3(cfdx::blocks_for(n_launch, 256)), dim3(256), 0, ctx->stream, (double *)U_C, (double *)Q_C,
                       (const int32_t *)Grid_P, det, (const double *)FLOW, (double *)ROE_AVER, (int)n_launch);
    return cfdx::check_launch("k_uq");
}

extern "C" int cfdx_viscous(cfdx_ctx *ctx, const cfdx_vec5 *U_C, cfdx_vec5 *F_C, cfdx_vec5 *E_C, cfdx_vec5 *G_C,
                            cfdx_vec5 *Fv_C, cfdx_vec5 *Ev_C, cfdx_vec5 *Gv_C, const cfdx_grid *Grid_P,
                            const cfdx_metric *metric, const double *det, cfdx_flow *FLOW, int64_t n_launch,
                            const cfdx_params *params)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (!U_C || !F_C || !E_C || !G_C || !Fv_C || !Ev_C || !Gv_C || !Grid_P || !metric || !det || !FLOW)
        return cfdx::fail(CFDX_E_INVALID, "cfdx_viscous: null buffer");
    if (n_launch < 0 || n_launch > INT32_MAX) return cfdx::fail(CFDX_E_INVALID, "cfdx_viscous: n_launch out of range");
    if (params && params->sgs != CFDX_SGS_MTS && params->sgs != CFDX_SGS_SMAGORINSKY)
        return cfdx::fail(CFDX_E_INVALID, "cfdx_viscous: sgs %d is neither CFDX_SGS_MTS nor CFDX_SGS_SMAGORINSKY", params->sgs);
    if (n_launch == 0) return CFDX_OK;
    hipLaunchKernelGGL(k_viscous, dim3(cfdx::blocks_for(n_launch, 128)), dim3(128), 0, ctx->stream, (const double *)U_C,
                       (double *)F_C, (double *)E_C, (double *)G_C, (double *)Fv_C, (double *)Ev_C, (double *)Gv_C,
                       (const int32_t *)Grid_P, (const double *)metric, det, (double *)FLOW, (int)n_launch, visc_params(params));
    return cfdx::check_launch("k_viscous");
}

extern "C" int cfdx_weno(cfdx_ctx *ctx, int j, double Mach, double del_t, const cfdx_deter *deter, const double *det,
                         cfdx_vec5 *U0, cfdx_vec5 *U1, cfdx_vec5 *U2, const cfdx_grid *Grid_P, const cfdx_vec5 *Q_C,
                         const cfdx_metric *metric, const cfdx_roe *ROE_AVER, const cfdx_vec5 *Fv, const cfdx_vec5 *Ev,
                         const cfdx_vec5 *Gv, const cfdx_vec5 *F, const cfdx_vec5 *E, const cfdx_vec5 *G, cfdx_flow *FLOW,
                         int64_t n_launch, const cfdx_params *params)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (!deter || !det || !U0 || !U1 || !U2 || !Grid_P || !Q_C || !metric || !ROE_AVER || !Fv || !Ev || !Gv || !F || !E || !G || !FLOW)
        return cfdx::fail(CFDX_E_INVALID, "cfdx_weno: null buffer");
    if (j < 0 || j > 2) return cfdx::fail(CFDX_E_INVALID, "cfdx_weno: stage j=%d not in 0..2", j);
    if (n_launch < 0 || n_launch > INT32_MAX) return cfdx::fail(CFDX_E_INVALID, "cfdx_weno: n_launch out of range");
    if (n_launch == 0) return CFDX_OK;
    cfdx_params d;
    if (!params) { cfdx_params_default(&d); params = &d; }
    hipLaunchKernelGGL(k_weno, dim3(cfdx::blocks_for(n_launch, 128)), dim3(128), 0, ctx->stream, j, Mach, del_t,
                       (const double *)deter, det, (double *)U0, (double *)U1, (double *)U2, (const int32_t *)Grid_P,
                       (const double *)Q_C, (const double *)metric, (const double *)ROE_AVER, (const double *)Fv,
                       (const double *)Ev, (const double *)Gv, (const double *)F, (const double *)E, (const double *)G,
                       (double *)FLOW, (int)n_launch, params->eps);
    return cfdx::check_launch("k_weno");
}
