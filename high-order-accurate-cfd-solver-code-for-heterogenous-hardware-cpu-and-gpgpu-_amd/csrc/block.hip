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

// Loads and stores at a 32-bit element index from a uniform base pointer, so the
// compiler emits SGPR-base + VGPR-offset global accesses (blocks hold < 2^29 nodes).
template <typename T>
__device__ __forceinline__ T ld(const T *base, uint32_t i)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st(T *base, uint32_t i, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + i * (uint32_t)sizeof(T)) = v;
}

struct UPtrs { const double *u[5]; };

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
    double *Fv[3][4];   // viscous fluxes, components 1..4 (the mass component is identically 0)
    double *mu;
    double *FH[3][5];
};

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b % 8),
// each XCD with its own L2. Remap the linear workgroup id so XCD x gets one contiguous
// range of tiles (a slab of k planes), keeping the +-1..3 plane neighbours of a tile in
// the same L2. Bijective for any count.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb)
{
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, l = b >> 3;
    return x * q + (x < r ? x : r) + l;
}

__device__ __forceinline__ uint32_t node_index(const Dims &D, int i, int j, int k) { return (uint32_t)i + (uint32_t)D.ni * ((uint32_t)j + (uint32_t)D.nj * (uint32_t)k); }

// index of the node at offset o along direction d from (i, j, k), periodic
__device__ __forceinline__ uint32_t shifted(const Dims &D, int i, int j, int k, int d, int o)
{
    if (d == 0) i = wrapi(i + o, D.ni);
    else if (d == 1) j = wrapi(j + o, D.nj);
    else k = wrapi(k + o, D.nk);
    return node_index(D, i, j, k);
}

// ---------------------------------------------------------------------------
// viscous node kernel (Viscous.cl:285-631 on a periodic block: central 4th order)
// ---------------------------------------------------------------------------
struct NodeTiles { uint32_t ntx, nty, nb; };

__device__ __forceinline__ void node_tile(const NodeTiles &T, int &i, int &j, int &k)
{
    const uint32_t t = xcd_tile(blockIdx.x, T.nb);
    const uint32_t tx = t % T.ntx, r = t / T.ntx, ty = r % T.nty;
    k = (int)(r / T.nty);
    i = (int)(tx * TX + threadIdx.x % TX);
    j = (int)(ty * TY + threadIdx.x / TX);
}

__global__ void __launch_bounds__(TX *TY) k_visc(Dims D, NodeTiles T, Geo G, State S, UPtrs Uin, ViscParams vp, int write_mu)
{
    int i, j, k;
    node_tile(T, i, j, k);
    if (i >= D.ni || j >= D.nj) return;
    const uint32_t g = node_index(D, i, j, k);
    double U[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) U[c] = ld(Uin.u[c], g);
    double dU[5][3], KS[3];
    const double iU0 = 1.0 / U[0];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const uint32_t m2 = shifted(D, i, j, k, d, -2), m1 = shifted(D, i, j, k, d, -1);
        const uint32_t p1 = shifted(D, i, j, k, d, 1), p2 = shifted(D, i, j, k, d, 2);
        double um1[5], up1[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const double *u = Uin.u[c];
            um1[c] = ld(u, m1);
            up1[c] = ld(u, p1);
            dU[c][d] = (1.0 / 12.0) * (8.0 * (up1[c] - um1[c]) - (ld(u, p2) - ld(u, m2)));
        }
        // MTS scale-similarity term of momentum component d+1 along d (Viscous.cl:493-495)
        const double a = (U[1 + d] * iU0) - ((0.25 * um1[1 + d] + 0.5 * U[1 + d] + 0.25 * up1[1 + d]) / (0.25 * um1[0] + 0.5 * U[0] + 0.25 * up1[0]));
        KS[d] = a * a;
    }
    double m[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) m[c] = ld(G.m[c], g);
    double fv[3][5];
    const double mu = viscous_node(dU, U, ld(S.P[1], g), ld(S.P[2], g), ld(S.P[3], g), m, ld(G.det, g), KS, vp, fv);
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int c = 1; c < 5; ++c) st(S.Fv[d][c - 1], g, fv[d][c]);
    if (write_mu) st(S.mu, g, mu);
}

// ---------------------------------------------------------------------------
// face kernel: F-hat at the +1/2 face of every node along direction DIR.
// A block takes a tile of TB faces along DIR times TA lines across it. Phase 1
// stages the conserved state Q and the node flux F_DIR of the TB+5 stencil
// points of every line in LDS (each point computed once per block); phase 2
// gives each thread one face, reading its six stencil points from LDS.
// ---------------------------------------------------------------------------
template <int DIR> struct FaceTile { static constexpr int TA = 16, TB = 16; };   // across = x, along = y or z
template <> struct FaceTile<0> { static constexpr int TA = 4, TB = 64; };      // along = x, across = y

template <int DIR>
__device__ __forceinline__ int lds_at(int c, int s, int a)
{
    constexpr int TA = FaceTile<DIR>::TA, NP = FaceTile<DIR>::TB + 5;
    return DIR == 0 ? (c * TA + a) * NP + s : (c * NP + s) * TA + a;
}

template <int DIR>
__device__ __forceinline__ uint32_t tile_node(const Dims &D, int along, int across, int fixed)
{
    if (DIR == 0) return node_index(D, along, across, fixed);
    if (DIR == 1) return node_index(D, across, along, fixed);
    return node_index(D, across, fixed, along);
}

template <int DIR>
__global__ void __launch_bounds__(256) k_face(Dims D, Geo G, State S, double eps)
{
    constexpr int TA = FaceTile<DIR>::TA, TB = FaceTile<DIR>::TB, NP = TB + 5;
    __shared__ double sq[5 * NP * TA];
    __shared__ double sf[5 * NP * TA];
    const int n_along = DIR == 0 ? D.ni : DIR == 1 ? D.nj : D.nk;
    const int n_across = DIR == 0 ? D.nj : D.ni;
    // tiles: DIR 0 (b-tile x, a-tile y, k), DIR 1 (a-tile x, b-tile y, k), DIR 2 (a-tile x, j, b-tile z)
    const uint32_t nta = (n_across + TA - 1) / TA, ntb = (n_along + TB - 1) / TB;
    const uint32_t nfix = DIR == 2 ? D.nj : D.nk;
    const uint32_t t = xcd_tile(blockIdx.x, nta * ntb * nfix);
    int b0, a0, fixed;
    if (DIR == 0) { b0 = (t % ntb) * TB; a0 = ((t / ntb) % nta) * TA; fixed = t / (ntb * nta); }
    else if (DIR == 1) { a0 = (t % nta) * TA; b0 = ((t / nta) % ntb) * TB; fixed = t / (nta * ntb); }
    else { a0 = (t % nta) * TA; fixed = (t / nta) % nfix; b0 = (t / (nta * nfix)) * TB; }
    const int tid = threadIdx.x;

    for (int p = tid; p < NP * TA; p += 256) {
        const int s = DIR == 0 ? p % NP : p / TA;
        const int a = DIR == 0 ? p / NP : p % TA;
        if (a0 + a >= n_across) continue;
        const int along = wrapi(b0 + s - 2, n_along);
        const uint32_t x = tile_node<DIR>(D, along, a0 + a, fixed);
        const double rho = ld(S.P[0], x), u = ld(S.P[1], x), v = ld(S.P[2], x), w = ld(S.P[3], x), pr = ld(S.P[4], x);
        double q[5], f[5];
        prim_to_q(rho, u, v, w, pr, q);
        node_flux(rho, u, v, w, pr, ld(G.m[3 * DIR], x), ld(G.m[3 * DIR + 1], x), ld(G.m[3 * DIR + 2], x), ld(G.det, x), f);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            sq[lds_at<DIR>(c, s, a)] = q[c];
            sf[lds_at<DIR>(c, s, a)] = f[c];
        }
    }
    __syncthreads();

    const int b = DIR == 0 ? tid % TB : tid / TA;
    const int a = DIR == 0 ? tid / TB : tid % TA;
    if (b0 + b >= n_along || a0 + a >= n_across) return;
    double q[6][5];
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int c = 0; c < 5; ++c) q[m][c] = sq[lds_at<DIR>(c, b + m, a)];
    const uint32_t gl = tile_node<DIR>(D, b0 + b, a0 + a, fixed);
    const uint32_t gr = tile_node<DIR>(D, wrapi(b0 + b + 1, n_along), a0 + a, fixed);
    double roe[5];
    {
        const double ra = ld(S.P[0], gl), ua = ld(S.P[1], gl), va = ld(S.P[2], gl), wa = ld(S.P[3], gl), pa = ld(S.P[4], gl);
        const double rb = ld(S.P[0], gr), ub = ld(S.P[1], gr), vb = ld(S.P[2], gr), wb = ld(S.P[3], gr), pb = ld(S.P[4], gr);
        roe_average(ra, ua, va, wa, enthalpy(ra, ua, va, wa, pa), rb, ub, vb, wb, enthalpy(rb, ub, vb, wb, pb), roe);
    }
    double llf[5];
    face_llf(q, roe, ld(G.kf[DIR][0], gl), ld(G.kf[DIR][1], gl), ld(G.kf[DIR][2], gl), ld(G.df[DIR], gl), eps, llf);
#pragma unroll
    for (int c = 0; c < 5; ++c)
        st(S.FH[DIR][c], gl, llf[c] + midpoint_correction(sf[lds_at<DIR>(c, b, a)], sf[lds_at<DIR>(c, b + 1, a)],
                                                          sf[lds_at<DIR>(c, b + 2, a)], sf[lds_at<DIR>(c, b + 3, a)],
                                                          sf[lds_at<DIR>(c, b + 4, a)], sf[lds_at<DIR>(c, b + 5, a)]));
}

template <int DIR>
dim3 face_grid(const Dims &D)
{
    constexpr int TA = FaceTile<DIR>::TA, TB = FaceTile<DIR>::TB;
    const int n_along = DIR == 0 ? D.ni : DIR == 1 ? D.nj : D.nk;
    const int n_across = DIR == 0 ? D.nj : D.ni;
    const int nfix = DIR == 2 ? D.nj : D.nk;
    return dim3((unsigned)(((n_across + TA - 1) / TA) * ((n_along + TB - 1) / TB) * nfix));
}

// ---------------------------------------------------------------------------
// update kernel: L and the SSP-RK3 stage (Weno.cl:3060-3135)
//   MODE 0..2: RK stage j; MODE 3: write L to Lout only
// ---------------------------------------------------------------------------
template <int MODE, bool VISC>
__global__ void __launch_bounds__(TX *TY) k_update(Dims D, NodeTiles T, Geo G, State S, double del_t, double *__restrict__ Lout)
{
    int i, j, k;
    node_tile(T, i, j, k);
    if (i >= D.ni || j >= D.nj) return;
    const uint32_t g = node_index(D, i, j, k);
    double L[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const uint32_t m1 = shifted(D, i, j, k, d, -1);
#pragma unroll
        for (int c = 0; c < 5; ++c) L[c] += ld(S.FH[d][c], g) - ld(S.FH[d][c], m1);
    }
    if (VISC) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const uint32_t m2 = shifted(D, i, j, k, d, -2), m1 = shifted(D, i, j, k, d, -1);
            const uint32_t p1 = shifted(D, i, j, k, d, 1), p2 = shifted(D, i, j, k, d, 2);
#pragma unroll
            for (int c = 1; c < 5; ++c) {
                const double *f = S.Fv[d][c - 1];
                L[c] += (1.0 / 12.0) * (8.0 * (ld(f, p1) - ld(f, m1)) - (ld(f, p2) - ld(f, m2)));
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
        for (int c = 0; c < 5; ++c) { un[c] = ld(S.U0[c], g) + del_t * L[c]; st(S.Us[c], g, un[c]); }
    } else if (MODE == 1) {
#pragma unroll
        for (int c = 0; c < 5; ++c) { un[c] = 0.75 * ld(S.U0[c], g) + 0.25 * ld(S.Us[c], g) + 0.25 * del_t * L[c]; st(S.Us[c], g, un[c]); }
    } else {
#pragma unroll
        for (int c = 0; c < 5; ++c) un[c] = (ld(S.U0[c], g) * (1.0 / 3.0)) + ((2.0 / 3.0) * ld(S.Us[c], g)) + (2.0 / 3.0) * (del_t * L[c]);
    }
    const double det = ld(G.det, g);
    const double idet = 1.0 / det;
    const double rho = idet * un[0], ir = 1.0 / rho;
    const double u = (idet * un[1]) * ir, v = (idet * un[2]) * ir, w = (idet * un[3]) * ir;
    const double e = (idet * un[4]) * ir - 0.5 * (u * u + v * v + w * w);
    const double p = e * rho * 0.4;
    st(S.P[0], g, rho); st(S.P[1], g, u); st(S.P[2], g, v); st(S.P[3], g, w); st(S.P[4], g, p);
    if (MODE == 2) {
        // next step's RK base: U_Q recomputes det*Q from the new primitives (Solver.cpp:377-388)
        double q[5];
        prim_to_q(rho, u, v, w, p, q);
#pragma unroll
        for (int c = 0; c < 5; ++c) st(S.U0[c], g, det * q[c]);
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

    Geo G{};
    State S{};
    double last_mach = 0.5;
    bool timing = false;
    hipEvent_t ev[8] = {};
    double t_visc = 0, t_face = 0, t_update = 0;
    int launches = 0;
};

namespace {

constexpr int ARRAYS = 22 + 5 + 5 + 5 + 12 + 1 + 15;

int check_params(const cfdx_params *p)
{
    if (!p) return cfdx::fail(CFDX_E_INVALID, "null params");
    if (p->precision != CFDX_FP64) return cfdx::fail(CFDX_E_UNSUPPORTED, "block: precision %d not built", p->precision);
    if (!p->viscous && !(std::isinf(p->reyl) && p->c_mts == 0.0))
        return cfdx::fail(CFDX_E_INVALID, "block: viscous=0 requires reyl=inf and c_mts=0 (else the RHS differs)");
    return CFDX_OK;
}

NodeTiles node_tiles(const Dims &D)
{
    NodeTiles T;
    T.ntx = (D.ni + TX - 1) / TX;
    T.nty = (D.nj + TY - 1) / TY;
    T.nb = T.ntx * T.nty * D.nk;
    return T;
}

void record(cfdx_block *b, int slot)
{
    if (b->timing) (void)hipEventRecord(b->ev[slot], b->ctx->stream);
}

// one stage; mode 3 = RHS only into Lout
int run_stage(cfdx_block *b, int mode, double del_t, const cfdx_params *p, double *Lout)
{
    const Dims &D = b->D;
    const NodeTiles T = node_tiles(D);
    const dim3 grid(T.nb), blk(TX * TY);
    hipStream_t st = b->ctx->stream;
    const bool visc = p->viscous != 0;
    const ViscParams vp = make_visc_params(p->mach, p->reyl, p->pr, p->free_t, p->c_mts, p->c_t);
    record(b, 0);
    if (visc) {
        // stage vector into the viscous kernel: U0 at j = 0, the raw U1 / U2 after (Solver.cpp:427-435)
        UPtrs uin;
        for (int c = 0; c < 5; ++c) uin.u[c] = (mode == 0 || mode == 3) ? b->S.U0[c] : b->S.Us[c];
        hipLaunchKernelGGL(k_visc, grid, blk, 0, st, D, T, b->G, b->S, uin, vp, (mode == 2 || mode == 3) ? 1 : 0);
        b->launches++;
    }
    record(b, 1);
    hipLaunchKernelGGL(k_face<0>, face_grid<0>(D), dim3(256), 0, st, D, b->G, b->S, p->eps);
    hipLaunchKernelGGL(k_face<1>, face_grid<1>(D), dim3(256), 0, st, D, b->G, b->S, p->eps);
    hipLaunchKernelGGL(k_face<2>, face_grid<2>(D), dim3(256), 0, st, D, b->G, b->S, p->eps);
    b->launches += 3;
    record(b, 2);
    if (visc) {
        switch (mode) {
        case 0: hipLaunchKernelGGL((k_update<0, true>), grid, blk, 0, st, D, T, b->G, b->S, del_t, Lout); break;
        case 1: hipLaunchKernelGGL((k_update<1, true>), grid, blk, 0, st, D, T, b->G, b->S, del_t, Lout); break;
        case 2: hipLaunchKernelGGL((k_update<2, true>), grid, blk, 0, st, D, T, b->G, b->S, del_t, Lout); break;
        default: hipLaunchKernelGGL((k_update<3, true>), grid, blk, 0, st, D, T, b->G, b->S, del_t, Lout); break;
        }
    } else {
        switch (mode) {
        case 0: hipLaunchKernelGGL((k_update<0, false>), grid, blk, 0, st, D, T, b->G, b->S, del_t, Lout); break;
        case 1: hipLaunchKernelGGL((k_update<1, false>), grid, blk, 0, st, D, T, b->G, b->S, del_t, Lout); break;
        case 2: hipLaunchKernelGGL((k_update<2, false>), grid, blk, 0, st, D, T, b->G, b->S, del_t, Lout); break;
        default: hipLaunchKernelGGL((k_update<3, false>), grid, blk, 0, st, D, T, b->G, b->S, del_t, Lout); break;
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
    if (n >= (1LL << 29)) return cfdx::fail(CFDX_E_CAPACITY, "block_create: %lld nodes, the limit is 2^29 (32-bit byte offsets)", n);
    cfdx_block *b = new cfdx_block;
    b->ctx = ctx;
    b->desc = *desc;
    b->D = Dims{desc->ni, desc->nj, desc->nk, n};
    hipError_t e = hipMalloc((void **)&b->mem, (size_t)ARRAYS * 8 * n);
    if (e != hipSuccess) { delete b; return cfdx::fail(CFDX_E_NOMEM, "block_create: hipMalloc %lld bytes: %s", (long long)ARRAYS * 8 * n, hipGetErrorString(e)); }
    double *p = b->mem;
    auto take = [&]() { double *r = p; p += n; return r; };
    b->G.det = take();
    for (auto &a : b->G.m) a = take();
    for (auto &r : b->G.kf) for (auto &a : r) a = take();
    for (auto &a : b->G.df) a = take();
    for (auto &a : b->S.P) a = take();
    for (auto &a : b->S.U0) a = take();
    for (auto &a : b->S.Us) a = take();
    for (auto &r : b->S.Fv) for (auto &a : r) a = take();   // 3 x 4
    b->S.mu = take();
    for (auto &r : b->S.FH) for (auto &a : r) a = take();
    e = hipMemset(b->mem, 0, (size_t)ARRAYS * 8 * n);
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
