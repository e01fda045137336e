// Structured-block fast path: one periodic ni x nj x nk block held in HBM as
// structure-of-arrays, stepped by
//   k_visc        node kernel: Fv, Ev, Gv (+ mu) from the stage vector U (Viscous.cl)
//   k_face<d>     one thread per face (i+1/2 along d): characteristic WCNS
//                 reconstruction, LLF and the 6th-order midpoint correction (Weno.cl)
//   k_update<j>   flux differences + viscous divergence -> L, SSP-RK3 stage j,
//                 primitive recovery (Weno.cl:3060-3135)
// Each face flux is computed once and shared by its two nodes (the reference
// computes it twice, once from each side, with identical operands).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "physics.hpp"

using namespace cfdx::dev;

namespace {

constexpr int TX = 64, TY = 4;

struct Dims {
    int ni, nj, nk;
    long long n;
};

// field arrays of the block (each n doubles, SoA)
struct Geo {
    const double *det;      // det at the node
    const double *m[9];     // node metrics zeta_x..xi_z
    const double *kf[3][3]; // face metric row of direction d at the +1/2 face (zeta_?ip, eta_?jp, xi_?kp)
    const double *df[3];    // half-point det at ip, jp, kp
};

struct State {
    double *P[5];   // rho, u, v, w, p
    double *U0[5];  // RK base det*Q
    double *Us[5];  // stage vector U1 / U2
    double *Fv[3][5];
    double *mu;
    double *FH[3][5];
};

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

__device__ __forceinline__ long long node_index(const Dims &D, int i, int j, int k) { return (long long)i + (long long)D.ni * (j + (long long)D.nj * k); }

// index of the node at offset o along direction d from (i, j, k), periodic
__device__ __forceinline__ long long shifted(const Dims &D, int i, int j, int k, int d, int o)
{
    if (d == 0) i = wrapi(i + o, D.ni);
    else if (d == 1) j = wrapi(j + o, D.nj);
    else k = wrapi(k + o, D.nk);
    return node_index(D, i, j, k);
}

// ---------------------------------------------------------------------------
// viscous node kernel (Viscous.cl:285-631 on a periodic block: central 4th order)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(TX *TY) k_visc(Dims D, Geo G, State S, const double *const *__restrict__ Uin, ViscParams vp)
{
    const int i = blockIdx.x * TX + threadIdx.x, j = blockIdx.y * TY + threadIdx.y, k = blockIdx.z;
    if (i >= D.ni || j >= D.nj) return;
    const long long g = node_index(D, i, j, k);
    double U[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) U[c] = Uin[c][g];
    double dU[5][3], KS[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long long m2 = shifted(D, i, j, k, d, -2), m1 = shifted(D, i, j, k, d, -1);
        const long long p1 = shifted(D, i, j, k, d, 1), p2 = shifted(D, i, j, k, d, 2);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const double *u = Uin[c];
            dU[c][d] = (1.0 / 12.0) * (8.0 * (u[p1] - u[m1]) - (u[p2] - u[m2]));
        }
        const double *u0 = Uin[0], *ud = Uin[1 + d];
        const double a = (U[1 + d] / U[0]) - ((0.25 * ud[m1] + 0.5 * U[1 + d] + 0.25 * ud[p1]) / (0.25 * u0[m1] + 0.5 * U[0] + 0.25 * u0[p1]));
        KS[d] = a * a;
    }
    double m[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) m[c] = G.m[c][g];
    double fv[3][5];
    const double mu = viscous_node(dU, U, S.P[1][g], S.P[2][g], S.P[3][g], m, G.det[g], KS, vp, fv);
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int c = 0; c < 5; ++c) S.Fv[d][c][g] = fv[d][c];
    S.mu[g] = mu;
}

// ---------------------------------------------------------------------------
// face kernel: F-hat at the +1/2 face of every node along direction DIR
// ---------------------------------------------------------------------------
template <int DIR>
__global__ void __launch_bounds__(TX *TY) k_face(Dims D, Geo G, State S, double eps)
{
    const int i = blockIdx.x * TX + threadIdx.x, j = blockIdx.y * TY + threadIdx.y, k = blockIdx.z;
    if (i >= D.ni || j >= D.nj) return;
    const long long g = node_index(D, i, j, k);
    double q[6][5], nf[6][5];
    double pl[5], pr[5];   // primitives of the two face nodes
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const long long x = shifted(D, i, j, k, DIR, s - 2);
        const double rho = S.P[0][x], u = S.P[1][x], v = S.P[2][x], w = S.P[3][x], p = S.P[4][x];
        prim_to_q(rho, u, v, w, p, q[s]);
        node_flux(rho, u, v, w, p, G.m[3 * DIR][x], G.m[3 * DIR + 1][x], G.m[3 * DIR + 2][x], G.det[x], nf[s]);
        if (s == 2) { pl[0] = rho; pl[1] = u; pl[2] = v; pl[3] = w; pl[4] = p; }
        if (s == 3) { pr[0] = rho; pr[1] = u; pr[2] = v; pr[3] = w; pr[4] = p; }
    }
    double roe[5];
    roe_average(pl[0], pl[1], pl[2], pl[3], enthalpy(pl[0], pl[1], pl[2], pl[3], pl[4]),
                pr[0], pr[1], pr[2], pr[3], enthalpy(pr[0], pr[1], pr[2], pr[3], pr[4]), roe);
    double out[5];
    face_flux(q, roe, G.kf[DIR][0][g], G.kf[DIR][1][g], G.kf[DIR][2][g], G.df[DIR][g], eps, nf, out);
#pragma unroll
    for (int c = 0; c < 5; ++c) S.FH[DIR][c][g] = out[c];
}

// ---------------------------------------------------------------------------
// update kernel: L and the SSP-RK3 stage (Weno.cl:3060-3135)
//   MODE 0..2: RK stage j; MODE 3: write L to Lout only
// ---------------------------------------------------------------------------
template <int MODE, bool VISC>
__global__ void __launch_bounds__(TX *TY) k_update(Dims D, Geo G, State S, double del_t, double *__restrict__ Lout)
{
    const int i = blockIdx.x * TX + threadIdx.x, j = blockIdx.y * TY + threadIdx.y, k = blockIdx.z;
    if (i >= D.ni || j >= D.nj) return;
    const long long g = node_index(D, i, j, k);
    double L[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long long m1 = shifted(D, i, j, k, d, -1);
#pragma unroll
        for (int c = 0; c < 5; ++c) L[c] += S.FH[d][c][g] - S.FH[d][c][m1];
    }
    if (VISC) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const long long m2 = shifted(D, i, j, k, d, -2), m1 = shifted(D, i, j, k, d, -1);
            const long long p1 = shifted(D, i, j, k, d, 1), p2 = shifted(D, i, j, k, d, 2);
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const double *f = S.Fv[d][c];
                L[c] += (1.0 / 12.0) * (8.0 * (f[p1] - f[m1]) - (f[p2] - f[m2]));
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) L[c] = -L[c];
    if (MODE == 3) {
#pragma unroll
        for (int c = 0; c < 5; ++c) Lout[(long long)c * D.n + g] = L[c];
        return;
    }
    double un[5];
    if (MODE == 0) {
#pragma unroll
        for (int c = 0; c < 5; ++c) { un[c] = S.U0[c][g] + del_t * L[c]; S.Us[c][g] = un[c]; }
    } else if (MODE == 1) {
#pragma unroll
        for (int c = 0; c < 5; ++c) { un[c] = 0.75 * S.U0[c][g] + 0.25 * S.Us[c][g] + 0.25 * del_t * L[c]; S.Us[c][g] = un[c]; }
    } else {
#pragma unroll
        for (int c = 0; c < 5; ++c) un[c] = (S.U0[c][g] / 3.0) + ((2.0 / 3.0) * S.Us[c][g]) + (2.0 / 3.0) * (del_t * L[c]);
    }
    const double det = G.det[g];
    const double idet = 1.0 / det;
    const double rho = idet * un[0], ir = 1.0 / rho;
    const double u = (idet * un[1]) * ir, v = (idet * un[2]) * ir, w = (idet * un[3]) * ir;
    const double e = (idet * un[4]) * ir - 0.5 * (u * u + v * v + w * w);
    const double p = e * rho * 0.4;
    S.P[0][g] = rho; S.P[1][g] = u; S.P[2][g] = v; S.P[3][g] = w; S.P[4][g] = p;
    if (MODE == 2) {
        // next step's RK base: U_Q recomputes det*Q from the new primitives (Solver.cpp:377-388)
        double q[5];
        prim_to_q(rho, u, v, w, p, q);
#pragma unroll
        for (int c = 0; c < 5; ++c) S.U0[c][g] = det * q[c];
    }
}

// U0 = det * Q(P) on upload
__global__ void k_base(Dims D, Geo G, State S)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= D.n) return;
    double q[5];
    prim_to_q(S.P[0][g], S.P[1][g], S.P[2][g], S.P[3][g], S.P[4][g], q);
    const double det = G.det[g];
#pragma unroll
    for (int c = 0; c < 5; ++c) S.U0[c][g] = det * q[c];
}

}  // namespace

// ---------------------------------------------------------------------------
// block object
// ---------------------------------------------------------------------------
struct cfdx_block {
    cfdx_ctx *ctx = nullptr;
    cfdx_block_desc desc{};
    Dims D{};
    double *mem = nullptr;      // one allocation, 68 arrays of n doubles
    double **uptrs = nullptr;   // device copy of the 3 stage-vector pointer sets (U0, Us)
    Geo G{};
    State S{};
    double last_mach = 0.5;
    bool timing = false;
    hipEvent_t ev[8] = {};
    double t_visc = 0, t_face = 0, t_update = 0;
    int launches = 0;
};

namespace {

constexpr int ARRAYS = 22 + 5 + 5 + 5 + 15 + 1 + 15;

int check_params(const cfdx_params *p)
{
    if (!p) return cfdx::fail(CFDX_E_INVALID, "null params");
    if (p->precision != CFDX_FP64) return cfdx::fail(CFDX_E_UNSUPPORTED, "block: precision %d not built", p->precision);
    if (!p->viscous && !(std::isinf(p->reyl) && p->c_mts == 0.0))
        return cfdx::fail(CFDX_E_INVALID, "block: viscous=0 requires reyl=inf and c_mts=0 (else the RHS differs)");
    return CFDX_OK;
}

dim3 grid_of(const Dims &D) { return dim3((D.ni + TX - 1) / TX, (D.nj + TY - 1) / TY, D.nk); }

void record(cfdx_block *b, int slot)
{
    if (b->timing) (void)hipEventRecord(b->ev[slot], b->ctx->stream);
}

// one stage; mode 3 = RHS only into Lout
int run_stage(cfdx_block *b, int mode, double del_t, const cfdx_params *p, double *Lout)
{
    const Dims &D = b->D;
    const dim3 grid = grid_of(D), blk(TX, TY);
    hipStream_t st = b->ctx->stream;
    const bool visc = p->viscous != 0;
    ViscParams vp{p->mach, p->reyl, p->pr, p->free_t, p->c_mts, p->c_t};
    record(b, 0);
    if (visc) {
        // stage vector into the viscous kernel: U0 at j = 0, the raw U1 / U2 after (Solver.cpp:427-435)
        const double *const *uin = (mode == 0 || mode == 3) ? (const double *const *)b->uptrs : (const double *const *)(b->uptrs + 5);
        hipLaunchKernelGGL(k_visc, grid, blk, 0, st, D, b->G, b->S, uin, vp);
        b->launches++;
    }
    record(b, 1);
    hipLaunchKernelGGL(k_face<0>, grid, blk, 0, st, D, b->G, b->S, p->eps);
    hipLaunchKernelGGL(k_face<1>, grid, blk, 0, st, D, b->G, b->S, p->eps);
    hipLaunchKernelGGL(k_face<2>, grid, blk, 0, st, D, b->G, b->S, p->eps);
    b->launches += 3;
    record(b, 2);
    if (visc) {
        switch (mode) {
        case 0: hipLaunchKernelGGL((k_update<0, true>), grid, blk, 0, st, D, b->G, b->S, del_t, Lout); break;
        case 1: hipLaunchKernelGGL((k_update<1, true>), grid, blk, 0, st, D, b->G, b->S, del_t, Lout); break;
        case 2: hipLaunchKernelGGL((k_update<2, true>), grid, blk, 0, st, D, b->G, b->S, del_t, Lout); break;
        default: hipLaunchKernelGGL((k_update<3, true>), grid, blk, 0, st, D, b->G, b->S, del_t, Lout); break;
        }
    } else {
        switch (mode) {
        case 0: hipLaunchKernelGGL((k_update<0, false>), grid, blk, 0, st, D, b->G, b->S, del_t, Lout); break;
        case 1: hipLaunchKernelGGL((k_update<1, false>), grid, blk, 0, st, D, b->G, b->S, del_t, Lout); break;
        case 2: hipLaunchKernelGGL((k_update<2, false>), grid, blk, 0, st, D, b->G, b->S, del_t, Lout); break;
        default: hipLaunchKernelGGL((k_update<3, false>), grid, blk, 0, st, D, b->G, b->S, del_t, Lout); break;
        }
    }
    b->launches++;
    record(b, 3);
    if (b->timing) {
        float a = 0, c = 0, e = 0;
        CFDX_HIP(hipEventSynchronize(b->ev[3]), "hipEventSynchronize");
        (void)hipEventElapsedTime(&a, b->ev[0], b->ev[1]);
        (void)hipEventElapsedTime(&c, b->ev[1], b->ev[2]);
        (void)hipEventElapsedTime(&e, b->ev[2], b->ev[3]);
        b->t_visc += a; b->t_face += c; b->t_update += e;
    }
    return cfdx::check_launch("block stage");
}

}  // namespace

extern "C" int64_t cfdx_block_bytes(const cfdx_block_desc *d)
{
    if (!d || d->ni < 1 || d->nj < 1 || d->nk < 1) return -1;
    return (int64_t)ARRAYS * 8 * (int64_t)d->ni * d->nj * d->nk + 10 * 8;
}

extern "C" int cfdx_block_create(cfdx_ctx *ctx, const cfdx_block_desc *desc, cfdx_block **out)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (!desc || !out) return cfdx::fail(CFDX_E_INVALID, "block_create: null argument");
    *out = nullptr;
    if (desc->ni < 3 || desc->nj < 3 || desc->nk < 3) return cfdx::fail(CFDX_E_INVALID, "block_create: each dimension must be >= 3");
    if (desc->periodic != 7) return cfdx::fail(CFDX_E_UNSUPPORTED, "block_create: only fully periodic blocks (periodic=7)");
    if (desc->precision != CFDX_FP64) return cfdx::fail(CFDX_E_UNSUPPORTED, "block_create: precision %d not built", desc->precision);
    const long long n = (long long)desc->ni * desc->nj * desc->nk;
    if (n > (1LL << 31)) return cfdx::fail(CFDX_E_CAPACITY, "block_create: %lld nodes exceed 2^31", n);
    cfdx_block *b = new cfdx_block;
    b->ctx = ctx;
    b->desc = *desc;
    b->D = Dims{desc->ni, desc->nj, desc->nk, n};
    hipError_t e = hipMalloc((void **)&b->mem, (size_t)ARRAYS * 8 * n);
    if (e != hipSuccess) { delete b; return cfdx::fail(CFDX_E_NOMEM, "block_create: hipMalloc %lld bytes: %s", (long long)ARRAYS * 8 * n, hipGetErrorString(e)); }
    e = hipMalloc((void **)&b->uptrs, 10 * sizeof(double *));
    if (e != hipSuccess) { (void)hipFree(b->mem); delete b; return cfdx::fail(CFDX_E_NOMEM, "block_create: hipMalloc"); }
    double *p = b->mem;
    auto take = [&]() { double *r = p; p += n; return r; };
    b->G.det = take();
    for (auto &a : b->G.m) a = take();
    for (auto &r : b->G.kf) for (auto &a : r) a = take();
    for (auto &a : b->G.df) a = take();
    for (auto &a : b->S.P) a = take();
    for (auto &a : b->S.U0) a = take();
    for (auto &a : b->S.Us) a = take();
    for (auto &r : b->S.Fv) for (auto &a : r) a = take();
    b->S.mu = take();
    for (auto &r : b->S.FH) for (auto &a : r) a = take();
    double *hp[10];
    for (int c = 0; c < 5; ++c) { hp[c] = b->S.U0[c]; hp[5 + c] = b->S.Us[c]; }
    e = hipMemcpy(b->uptrs, hp, sizeof hp, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(b->mem, 0, (size_t)ARRAYS * 8 * n);
    for (auto &ev : b->ev) if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e != hipSuccess) { cfdx_block_destroy(b); return cfdx::hip_fail(e, "block_create"); }
    *out = b;
    return CFDX_OK;
}

extern "C" int cfdx_block_destroy(cfdx_block *b)
{
    if (!b) return CFDX_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (auto &ev : b->ev) if (ev) (void)hipEventDestroy(ev);
    if (b->mem) (void)hipFree(b->mem);
    if (b->uptrs) (void)hipFree(b->uptrs);
    delete b;
    return CFDX_OK;
}

extern "C" int cfdx_block_set_geometry(cfdx_block *b, const double *det, const cfdx_metric *metric, const cfdx_deter *deter)
{
    if (!b || !det || !metric || !deter) return cfdx::fail(CFDX_E_INVALID, "set_geometry: null argument");
    CFDX_TRY(cfdx::enter(b->ctx));
    const Dims &D = b->D;
    const long long n = D.n;
    // face-consistency: the minus face of node i must be the plus face of its -d neighbour
    // (Main.cpp:2368-2435 builds both from the same six node values)
    for (int k = 0; k < D.nk; ++k)
        for (int j = 0; j < D.nj; ++j)
            for (int i = 0; i < D.ni; ++i) {
                const long long g = i + (long long)D.ni * (j + (long long)D.nj * k);
                const long long nb[3] = {((i + D.ni - 1) % D.ni) + (long long)D.ni * (j + (long long)D.nj * k),
                                         i + (long long)D.ni * (((j + D.nj - 1) % D.nj) + (long long)D.nj * k),
                                         i + (long long)D.ni * (j + (long long)D.nj * ((k + D.nk - 1) % D.nk))};
                const double *dp = &deter[g].ip;
                const double *dn = &deter[nb[0]].ip;
                for (int d = 0; d < 3; ++d) {
                    const double *dm = &deter[nb[d]].ip;
                    const double a = dp[2 * d + 1], c = dm[2 * d];
                    if (std::fabs(a - c) > 1e-13 * std::fabs(c)) return cfdx::fail(CFDX_E_INVALID, "set_geometry: deter of node %lld dir %d is not face-consistent", g, d);
                    for (int q = 0; q < 3; ++q) {
                        const double x = metric[g].face[2 * d + 1][3 * d + q], y = metric[nb[d]].face[2 * d][3 * d + q];
                        if (std::fabs(x - y) > 1e-13 * (std::fabs(y) + 1e-300) && std::fabs(x - y) > 1e-15)
                            return cfdx::fail(CFDX_E_INVALID, "set_geometry: face metric of node %lld dir %d is not face-consistent", g, d);
                    }
                }
                (void)dn;
            }
    std::vector<double> h((size_t)22 * n);
    for (long long g = 0; g < n; ++g) {
        h[g] = det[g];
        for (int c = 0; c < 9; ++c) h[(1 + c) * n + g] = metric[g].node[c];
        for (int d = 0; d < 3; ++d)
            for (int c = 0; c < 3; ++c) h[(10 + 3 * d + c) * n + g] = metric[g].face[2 * d][3 * d + c];
        const double *dd = &deter[g].ip;
        for (int d = 0; d < 3; ++d) h[(19 + d) * n + g] = dd[2 * d];
    }
    CFDX_HIP(hipMemcpyAsync((void *)b->G.det, h.data(), h.size() * 8, hipMemcpyHostToDevice, b->ctx->stream), "set_geometry copy");
    CFDX_HIP(hipStreamSynchronize(b->ctx->stream), "set_geometry sync");
    return CFDX_OK;
}

extern "C" int cfdx_block_set_flow(cfdx_block *b, const cfdx_flow *flow)
{
    if (!b || !flow) return cfdx::fail(CFDX_E_INVALID, "set_flow: null argument");
    CFDX_TRY(cfdx::enter(b->ctx));
    const long long n = b->D.n;
    std::vector<double> h((size_t)5 * n);
    for (long long g = 0; g < n; ++g) {
        h[g] = flow[g].rho; h[n + g] = flow[g].u; h[2 * n + g] = flow[g].v; h[3 * n + g] = flow[g].w; h[4 * n + g] = flow[g].p;
    }
    CFDX_HIP(hipMemcpyAsync(b->S.P[0], h.data(), h.size() * 8, hipMemcpyHostToDevice, b->ctx->stream), "set_flow copy");
    hipLaunchKernelGGL(k_base, dim3(cfdx::blocks_for(n, 256)), dim3(256), 0, b->ctx->stream, b->D, b->G, b->S);
    CFDX_TRY(cfdx::check_launch("k_base"));
    CFDX_HIP(hipStreamSynchronize(b->ctx->stream), "set_flow sync");
    return CFDX_OK;
}

extern "C" int cfdx_block_get_flow(cfdx_block *b, cfdx_flow *flow)
{
    if (!b || !flow) return cfdx::fail(CFDX_E_INVALID, "get_flow: null argument");
    CFDX_TRY(cfdx::enter(b->ctx));
    const long long n = b->D.n;
    std::vector<double> h((size_t)6 * n);
    CFDX_HIP(hipMemcpyAsync(h.data(), b->S.P[0], (size_t)5 * n * 8, hipMemcpyDeviceToHost, b->ctx->stream), "get_flow copy");
    CFDX_HIP(hipMemcpyAsync(h.data() + 5 * n, b->S.mu, (size_t)n * 8, hipMemcpyDeviceToHost, b->ctx->stream), "get_flow copy");
    CFDX_HIP(hipStreamSynchronize(b->ctx->stream), "get_flow sync");
    const double M = b->last_mach;
    for (long long g = 0; g < n; ++g) {
        cfdx_flow &f = flow[g];
        f.rho = h[g]; f.u = h[n + g]; f.v = h[2 * n + g]; f.w = h[3 * n + g]; f.p = h[4 * n + g]; f.mu = h[5 * n + g];
        // Weno.cl:3132-3135
        f.e = f.p / (0.4 * f.rho);
        f.t = (1.4 * M * M) * (f.p / f.rho);
        f.a = std::sqrt(1.4 * f.p / f.rho);
    }
    return CFDX_OK;
}

extern "C" int cfdx_block_step(cfdx_block *b, double Mach, double del_t, int nsteps, const cfdx_params *params)
{
    if (!b) return cfdx::fail(CFDX_E_INVALID, "block_step: null block");
    CFDX_TRY(check_params(params));
    CFDX_TRY(cfdx::enter(b->ctx));
    if (nsteps < 0) return cfdx::fail(CFDX_E_INVALID, "block_step: nsteps < 0");
    b->last_mach = Mach;
    b->t_visc = b->t_face = b->t_update = 0;
    b->launches = 0;
    for (int s = 0; s < nsteps; ++s)
        for (int j = 0; j < 3; ++j) CFDX_TRY(run_stage(b, j, del_t, params, nullptr));
    return CFDX_OK;
}

extern "C" int cfdx_block_rhs(cfdx_block *b, const cfdx_params *params, double *L_host)
{
    if (!b || !L_host) return cfdx::fail(CFDX_E_INVALID, "block_rhs: null argument");
    CFDX_TRY(check_params(params));
    CFDX_TRY(cfdx::enter(b->ctx));
    const long long n = b->D.n;
    double *Ld = nullptr;
    CFDX_HIP(hipMalloc((void **)&Ld, (size_t)5 * n * 8), "block_rhs hipMalloc");
    int rc = run_stage(b, 3, 0.0, params, Ld);
    if (rc == CFDX_OK) {
        hipError_t e = hipMemcpyAsync(L_host, Ld, (size_t)5 * n * 8, hipMemcpyDeviceToHost, b->ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->ctx->stream);
        if (e != hipSuccess) rc = cfdx::hip_fail(e, "block_rhs copy");
    }
    (void)hipFree(Ld);
    return rc;
}

extern "C" int cfdx_block_set_timing(cfdx_block *b, int on)
{
    if (!b) return cfdx::fail(CFDX_E_INVALID, "set_timing: null block");
    b->timing = on != 0;
    return CFDX_OK;
}

extern "C" int cfdx_block_timing(cfdx_block *b, double *ms_visc, double *ms_face, double *ms_update, int *launches)
{
    if (!b) return cfdx::fail(CFDX_E_INVALID, "timing: null block");
    if (ms_visc) *ms_visc = b->t_visc;
    if (ms_face) *ms_face = b->t_face;
    if (ms_update) *ms_update = b->t_update;
    if (launches) *launches = b->launches;
    return CFDX_OK;
}
