// Structured-block kernels. Each block is one periodic ni x nj x nk grid held in HBM as
// structure-of-arrays; a launch covers a list of blocks (one for cfdx_block_step, many for a
// batch), every workgroup taking one tile of one block through a prefix table of tile counts.
//   k_visc        node kernel: Fv, Ev, Gv (+ mu) from the stage vector U (Viscous.cl)
//   k_face<DIR>   one thread per face (i+1/2 along DIR): characteristic WCNS reconstruction,
//                 LLF and the 6th-order midpoint correction (Weno.cl), LDS-staged tiles
//   k_update      flux differences + viscous divergence -> L, SSP-RK3 stage j, primitives
//                 (Weno.cl:3060-3135)
// Each face flux is computed once and shared by its two nodes (the reference computes it
// twice, once from each side, with identical operands).
// S is the flux arithmetic / storage type: double (plain form) or float (mixed form). The RK
// base U0, the stage vector Us, the metrics and the L accumulation are double in both.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "block.hpp"

using namespace cfdx::dev;

namespace cfdx {
namespace blk {

// Loads and stores at a 32-bit element index from a uniform base pointer, so the compiler
// emits SGPR-base + VGPR-offset accesses (a block holds < 2^29 nodes). The array pointers come
// from a descriptor in memory, so they are cast to the global address space explicitly: global
// (not flat) instructions count only on the vector-memory wait counter, not on the LDS one.
#define CFDX_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T ld(const T *base, uint32_t i)
{
    const CFDX_GLOBAL char *p = (const CFDX_GLOBAL char *)base;
    return *(const CFDX_GLOBAL T *)(p + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st(T *base, uint32_t i, T v)
{
    CFDX_GLOBAL char *p = (CFDX_GLOBAL char *)base;
    *(CFDX_GLOBAL T *)(p + i * (uint32_t)sizeof(T)) = v;
}

// Streaming forms for data a kernel touches once (geometry read at its own node, outputs): the
// non-temporal hint lets the L2 evict these lines first, keeping the stencil neighbours that
// adjacent tiles read again. Used for periodic blocks only (measured at 128^3: the whole step
// 5 % faster in FP64 and mixed, 64 cubes 4 %); on the bounded C-grid (384x128x32) the hint
// costs 1 % in FP64 and 4 % mixed, so lists with bounded blocks keep plain accesses.
template <bool NT, typename T>
__device__ __forceinline__ T ldx(const T *base, uint32_t i)
{
    if constexpr (NT)
        return __builtin_nontemporal_load((const CFDX_GLOBAL T *)((const CFDX_GLOBAL char *)base + i * (uint32_t)sizeof(T)));
    else
        return ld(base, i);
}
template <bool NT, typename T>
__device__ __forceinline__ void stx(T *base, uint32_t i, T v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, (CFDX_GLOBAL T *)((CFDX_GLOBAL char *)base + i * (uint32_t)sizeof(T)));
    else
        st(base, i, v);
}

// Geometry array a (0 det, 1-9 node metrics, 10-18 face metric rows, 19-21 face dets) in the flux
// type: the double arrays in an FP64 block, their float copies in a mixed block (the flux
// arithmetic reads the metrics in S; the copies hold the same roundings a conversion at the load
// would give, in half the bytes).
template <typename S>
__device__ __forceinline__ const S *geo(const BlockDev &B, int a);
template <>
__device__ __forceinline__ const double *geo<double>(const BlockDev &B, int a) { return B.det + (size_t)a * B.n; }
template <>
__device__ __forceinline__ const float *geo<float>(const BlockDev &B, int a) { return B.geo32 + (size_t)a * B.n; }

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// stored index of owned coordinates (i, j, k); a bounded direction may go 3 nodes beyond the
// owned range into its ghost layers
__device__ __forceinline__ uint32_t node_index(const BlockDev &B, int i, int j, int k)
{
    return (uint32_t)(i + B.off[0]) + (uint32_t)B.pi * ((uint32_t)(j + B.off[1]) + (uint32_t)B.pj * (uint32_t)(k + B.off[2]));
}

// owned coordinate c + o along a direction of n owned nodes: wrapped if the direction is periodic
__device__ __forceinline__ int along_shift(const BlockDev &B, int d, int c, int o, int n)
{
    return B.off[d] ? c + o : wrapi(c + o, n);
}

__device__ __forceinline__ uint32_t shifted(const BlockDev &B, int i, int j, int k, int d, int o)
{
    if (d == 0) i = along_shift(B, 0, i, o, B.ni);
    else if (d == 1) j = along_shift(B, 1, j, o, B.nj);
    else k = along_shift(B, 2, k, o, B.nk);
    return node_index(B, i, j, k);
}

// number of +1/2 faces a sweep along d computes per line: one more than the owned nodes in a
// bounded direction (the face between the first owned node and its ghost)
__device__ __host__ __forceinline__ int faces_along(const BlockDev &B, int d)
{
    const int n = d == 0 ? B.ni : d == 1 ? B.nj : B.nk;
    return B.off[d] ? n + 1 : n;
}

// The loc-selected forms of a bounded block reduce to the central ones unless the node is a side
// node (code 1) or a side node or ghost (code 100) lies within 2 nodes of it along d; elsewhere
// the kernels take the central forms directly, so a wave away from the sides keeps the short path.
__device__ __forceinline__ bool near_side(const BlockDev &B, int i, int j, int k, int d)
{
    const int c[3] = {i, j, k}, n[3] = {B.ni, B.nj, B.nk};
    bool s = false;
#pragma unroll
    for (int q = 0; q < 3; ++q) s |= B.off[q] && (c[q] == 0 || c[q] == n[q] - 1);
    return s || (B.off[d] && (c[d] < (int)SHELL || c[d] > n[d] - 1 - (int)SHELL));
}

// The shell of a bounded block: the nodes within SHELL of a side along a bounded direction (the
// only nodes where a loc-selected form can differ from the central one). The node kernels skip it and
// the shell kernels cover it, so the tiled kernels carry no loc branches. Enumeration: the k shell
// planes (if k is bounded), then the j shell rows inside the k range left, then the i shell
// columns inside the j and k ranges left; i runs fastest in the first two parts.
__device__ __forceinline__ bool in_shell(const BlockDev &B, int i, int j, int k)
{
    const int c[3] = {i, j, k}, n[3] = {B.ni, B.nj, B.nk};
    bool s = false;
#pragma unroll
    for (int q = 0; q < 3; ++q) s |= B.off[q] && (c[q] < (int)SHELL || c[q] > n[q] - 1 - (int)SHELL);
    return s;
}

__device__ __forceinline__ void shell_node(const BlockDev &B, uint32_t t, int &i, int &j, int &k)
{
    const uint32_t ni = B.ni, nj = B.nj, nk = B.nk;
    const uint32_t kb = B.off[2] ? SHELL : 0u, jb = B.off[1] ? SHELL : 0u, w = 2 * SHELL;
    auto layer = [](uint32_t m, uint32_t n) { return (int)(m < SHELL ? m : n - 2 * SHELL + m); };
    if (B.off[2]) {
        const uint32_t part = w * ni * nj;
        if (t < part) {
            i = (int)(t % ni);
            j = (int)((t / ni) % nj);
            k = layer(t / (ni * nj), nk);
            return;
        }
        t -= part;
    }
    if (B.off[1]) {
        const uint32_t part = w * ni * (nk - 2 * kb);
        if (t < part) {
            i = (int)(t % ni);
            j = layer((t / ni) % w, nj);
            k = (int)(kb + t / (w * ni));
            return;
        }
        t -= part;
    }
    i = layer(t % w, ni);
    j = (int)(jb + (t / w) % (nj - 2 * jb));
    k = (int)(kb + t / (w * (nj - 2 * jb)));
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b % 8), each
// with its own L2. Remap the linear workgroup id so XCD x takes one contiguous range of tiles
// (a slab of k planes of one block). Bijective for any count.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb)
{
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, l = b >> 3;
    return x * q + (x < r ? x : r) + l;
}

// Node tile lt of a block: (i0, j0) of its first node and its plane k. Tiles run along i, then
// (with B.strip = S > 0) through strips of S tile rows, each strip plane by plane, so that the
// k +- 2 planes of a strip are recent in the L2 of the XCD that walks it; S = 0: plane by plane.
__device__ __forceinline__ void node_tile(const BlockDev &B, uint32_t lt, int TX, int TY, int &i0, int &j0, int &k)
{
    const uint32_t ntx = (B.ni + TX - 1) / TX, nty = (B.nj + TY - 1) / TY, nk = (uint32_t)B.nk;
    const uint32_t S = B.strip;
    if (S == 0 || S >= nty) {
        i0 = (int)((lt % ntx) * TX);
        j0 = (int)(((lt / ntx) % nty) * TY);
        k = (int)(lt / (ntx * nty));
        return;
    }
    const uint32_t per_full = ntx * S * nk, full = nty / S;
    uint32_t sidx, r, rows;
    if (lt < full * per_full) { sidx = lt / per_full; r = lt % per_full; rows = S; }
    else { sidx = full; r = lt - full * per_full; rows = nty - full * S; }
    k = (int)(r / (ntx * rows));
    const uint32_t q = r % (ntx * rows);
    i0 = (int)((q % ntx) * TX);
    j0 = (int)((sidx * S + q / ntx) * TY);
}

// block of a global tile id: the largest b with prefix[b] <= t (prefix[0] = 0, prefix[nblk] = total)
__device__ __forceinline__ uint32_t find_block(const uint32_t *__restrict__ prefix, uint32_t nblk, uint32_t t)
{
    uint32_t lo = 0, hi = nblk;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// viscous node kernel (Viscous.cl:285-631 on a periodic block: central 4th order)
// ---------------------------------------------------------------------------
// Fv, Ev, Gv (+ mu) of one node. LOC: a shell node of a bounded block (loc-selected forms near
// the sides); otherwise the central forms.
// PT: the stage vector of the node's in-plane stencil (i and j offsets -2 .. 2) is read from a
// plane tile in LDS: pt[(c * rows + ty + 2 + dj) * cols + tx + 2 + di] (the same doubles the global
// loads would give); the k offsets are read from global memory.
template <typename S, bool LOC, bool NT, bool PT = false>
__device__ __forceinline__ void visc_node(const BlockDev &B, int i, int j, int k, int mode, const double *pt = nullptr,
                                          int tx = 0, int ty = 0, int cols = 0, int rows = 0)
{
    const uint32_t g = node_index(B, i, j, k);
    // stage vector into the viscous kernel: U0 at j = 0, the raw U1 / U2 after (Solver.cpp:427-435)
    const double *const *Uin = (mode == 0 || mode == 3) ? B.U0 : B.Us;
    auto tile = [&](int c, int d, int o) -> double {
        return pt[(c * rows + ty + 2 + (d == 1 ? o : 0)) * cols + tx + 2 + (d == 0 ? o : 0)];
    };
    double U[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) U[c] = PT ? tile(c, 0, 0) : ld(Uin[c], g);
    double dU[5][3], KS[3];
    const double iU0 = pinv_(U[0]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const uint32_t m2 = shifted(B, i, j, k, d, -2), m1 = shifted(B, i, j, k, d, -1);
        const uint32_t p1 = shifted(B, i, j, k, d, 1), p2 = shifted(B, i, j, k, d, 2);
        double um1[5], up1[5];
        if (LOC && near_side(B, i, j, k, d)) {
            // bounded block: the loc-selected one-sided forms of Viscous.cl:285-435 near the sides
            const uint32_t m3 = shifted(B, i, j, k, d, -3), p3 = shifted(B, i, j, k, d, 3);
            const int LI = ld(B.code, g), LM2 = ld(B.code, m2), LM1 = ld(B.code, m1), LP1 = ld(B.code, p1),
                      LP2 = ld(B.code, p2);
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const double *u = Uin[c];
                um1[c] = ld(u, m1);
                up1[c] = ld(u, p1);
                const double f[7] = {ld(u, m3), ld(u, m2), um1[c], U[c], up1[c], ld(u, p2), ld(u, p3)};
                dU[c][d] = loc_derivative(f, LI, LM2, LM1, LP1, LP2, false);
            }
        } else if (PT && d < 2) {
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                um1[c] = tile(c, d, -1);
                up1[c] = tile(c, d, 1);
                dU[c][d] = (1.0 / 12.0) * (8.0 * (up1[c] - um1[c]) - (tile(c, d, 2) - tile(c, d, -2)));
            }
        } else {
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const double *u = Uin[c];
                um1[c] = ld(u, m1);
                up1[c] = ld(u, p1);
                dU[c][d] = (1.0 / 12.0) * (8.0 * (up1[c] - um1[c]) - (ld(u, p2) - ld(u, m2)));
            }
        }
        // MTS scale-similarity term of momentum component d+1 along d (Viscous.cl:493-495)
        const double a = (U[1 + d] * iU0) - ((0.25 * um1[1 + d] + 0.5 * U[1 + d] + 0.25 * up1[1 + d]) * pinv_(0.25 * um1[0] + 0.5 * U[0] + 0.25 * up1[0]));
        KS[d] = a * a;
    }
    // the differences above are taken in double; the flux arithmetic runs in S
    S dUs[5][3], Uv[5], KSs[3], m[9];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        Uv[c] = S(U[c]);
#pragma unroll
        for (int d = 0; d < 3; ++d) dUs[c][d] = S(dU[c][d]);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) KSs[d] = S(KS[d]);
#pragma unroll
    for (int c = 0; c < 9; ++c) m[c] = ldx<NT>(geo<S>(B, 1 + c), g);
    S fv[3][5];
    // the node velocity: away from the sides, from the stage vector already in registers (U_m /
    // U_0, the primitives of P up to rounding) instead of three more loads; near the sides the
    // side conditions set P apart from the stage vector, so it is read from P
    S vel[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) vel[c] = LOC ? ld((const S *)B.P[1 + c], g) : S(U[1 + c] * iU0);
    const S mu = viscous_node(dUs, Uv, vel[0], vel[1], vel[2], m,
                              ldx<NT>(geo<S>(B, 0), g), KSs, B.vp, fv);
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int c = 1; c < 5; ++c) stx<NT>((S *)B.Fv[d][c - 1], g, fv[d][c]);
    if (mode >= 2) stx<NT>(B.mu, g, (double)mu);
}

// Node tiles of the list first (tile ids through the XCD remap), then, for lists with bounded
// blocks, nblk x shell_tiles shell tiles of TX * TY shell nodes each (block = id / shell_tiles),
// dispatched last in plain order so that they spread over the XCDs at the end of the launch.
// Tiles of at least 4 rows stage the in-plane stencil states of the tile in LDS (PLANE): each
// state is loaded once per tile instead of once per node that reaches it.
template <typename S, int TX, int TY, bool BND>
__global__ void __launch_bounds__(TX *TY) k_visc(const BlockDev *__restrict__ blks, const uint32_t *__restrict__ prefix,
                                                           uint32_t nblk, uint32_t total, int mode, uint32_t shell_tiles)
{
    constexpr bool PLANE = TY >= 4;
    constexpr int PC = TX + 4, PR = TY + 4;
    __shared__ double plane[PLANE ? 5 * PR * PC : 1];
    const uint32_t all = total + (BND ? nblk * shell_tiles : 0u);
    for (uint32_t wg = blockIdx.x; wg < all; wg += gridDim.x) {
        if (BND && wg >= total) {
            const uint32_t s = wg - total;
            const BlockDev &B = blks[s / shell_tiles];
            const uint32_t t = (s % shell_tiles) * (TX * TY) + threadIdx.x;
            if (!B.viscous || t >= shell_count(B)) continue;
            int i, j, k;
            shell_node(B, t, i, j, k);
            visc_node<S, true, !BND>(B, i, j, k, mode);
            continue;
        }
        const uint32_t t = xcd_tile(wg, total);
        const BlockDev &B = blks[find_block(prefix, nblk, t)];
        if (!B.viscous) continue;
        const uint32_t lt = t - B.tile0[0];
        int i0, j0, k;
        node_tile(B, lt, TX, TY, i0, j0, k);
        const int tx = (int)(threadIdx.x % TX), ty = (int)(threadIdx.x / TX);
        const int i = i0 + tx, j = j0 + ty;
        if constexpr (PLANE) {
            // the stage vector of the tile and its in-plane halo (2 nodes on each side, corners
            // left out) into LDS; points past the last node a tile node reaches are skipped
            const double *const *Uin = (mode == 0 || mode == 3) ? B.U0 : B.Us;
            __syncthreads();   // the previous tile's plane fully read
            for (int p = (int)threadIdx.x; p < PR * PC; p += TX * TY) {
                const int r = p / PC, q = p % PC;
                const bool hi = q >= 2 && q < TX + 2, hj = r >= 2 && r < TY + 2;
                if (!hi && !hj) continue;
                if (i0 + q - 2 > B.ni + 1 || j0 + r - 2 > B.nj + 1) continue;
                const uint32_t x = node_index(B, along_shift(B, 0, i0, q - 2, B.ni), along_shift(B, 1, j0, r - 2, B.nj), k);
#pragma unroll
                for (int c = 0; c < 5; ++c) plane[(c * PR + r) * PC + q] = ld(Uin[c], x);
            }
            __syncthreads();
        }
        if (i >= B.ni || j >= B.nj) continue;
        if (BND && B.code && in_shell(B, i, j, k)) continue;   // a shell tile's node
        if constexpr (PLANE) visc_node<S, false, !BND, true>(B, i, j, k, mode, plane, tx, ty, PC, PR);
        else visc_node<S, false, !BND>(B, i, j, k, mode);
    }
}

// ---------------------------------------------------------------------------
// face kernel: F-hat at the +1/2 face of every node along DIR. A tile is TB faces along DIR
// times TA lines across it. Phase 1 stages the conserved state Q and the node flux F_DIR of the
// TB+5 stencil points of every line in LDS (each point computed once per tile); phase 2 gives
// each thread one face, reading its six stencil points from LDS.
// ---------------------------------------------------------------------------
template <int DIR, int TA, int TB>
__device__ __forceinline__ int lds_at(int c, int s, int a)
{
    constexpr int NP = TB + 5;
    return DIR == 0 ? (c * TA + a) * NP + s : (c * NP + s) * TA + a;
}

template <int DIR>
__device__ __forceinline__ uint32_t tile_node(const BlockDev &B, int along, int across, int fixed)
{
    if (DIR == 0) return node_index(B, along, across, fixed);
    if (DIR == 1) return node_index(B, across, along, fixed);
    return node_index(B, across, fixed, along);
}

// Waves per SIMD the face kernel is compiled for: 4 (at most 128 VGPRs) for the tiles whose FP64
// form fits them without spilling (every i tile, the (32, 8) tile along j and k), else 3. A
// kernel that spills needs scratch memory, and its launches ran up to 40 % slower in a sequence of
// sweeps than at 3 waves without.
template <int DIR, int TA, int TB>
constexpr int FACE_WAVES = (DIR == 0 || (TA == 32 && TB == 8)) ? 4 : 3;

template <typename S, int DIR, int TA, int TB>
__global__ void __launch_bounds__(256, (FACE_WAVES<DIR, TA, TB>)) k_face(const BlockDev *__restrict__ blks, const uint32_t *__restrict__ prefix,
                                              uint32_t nblk, uint32_t total)
{
    static_assert(TA * TB == 256, "face tiles have 256 threads");
    constexpr int NP = TB + 5;
    // mixed form: the Roe state's inputs (rho, u, v, w, h) of the stencil points are staged too, so
    // phase 2 reads the two centre nodes from LDS; FP64 derives them from the conserved states of
    // the two centre nodes, which phase 2 holds in registers anyway (no primitive loads; staging
    // them in FP64 would cost a wave per SIMD).
    constexpr bool STAGE_ROE = std::is_same<S, float>::value;
    __shared__ S sq[5 * NP * TA];
    __shared__ S sf[5 * NP * TA];
    __shared__ S sp[STAGE_ROE ? 5 * NP * TA : 1];
    const int tid = threadIdx.x;
    for (uint32_t wg = blockIdx.x; wg < total; wg += gridDim.x) {
        const uint32_t t = xcd_tile(wg, total);
        const BlockDev &B = blks[find_block(prefix, nblk, t)];
        const uint32_t lt = t - B.tile0[1 + DIR];
        const int n_along = DIR == 0 ? B.ni : DIR == 1 ? B.nj : B.nk;
        const int n_across = DIR == 0 ? B.nj : B.ni;
        // face f of a line is the +1/2 face of along-coordinate f - sh (sh = 1 in a bounded
        // direction: the first face lies between the first owned node and its ghost)
        const int nfa = faces_along(B, DIR), sh = B.off[DIR] ? 1 : 0;
        const uint32_t nta = (n_across + TA - 1) / TA, ntb = (nfa + TB - 1) / TB;
        // tiles run along the lines' direction before the fixed coordinate, so that the next tile of
        // an XCD re-reads the stencil halo planes of this one from its L2
        int b0, a0, fixed;
        if (DIR == 0) { b0 = (lt % ntb) * TB; a0 = ((lt / ntb) % nta) * TA; fixed = lt / (ntb * nta); }
        else { a0 = (lt % nta) * TA; b0 = ((lt / nta) % ntb) * TB; fixed = lt / (nta * ntb); }
        const S *P0 = (const S *)B.P[0], *P1 = (const S *)B.P[1], *P2 = (const S *)B.P[2], *P3 = (const S *)B.P[3], *P4 = (const S *)B.P[4];

        __syncthreads();   // the previous tile's LDS fully consumed
        for (int p = tid; p < NP * TA; p += 256) {
            const int s = DIR == 0 ? p % NP : p / TA;
            const int a = DIR == 0 ? p / NP : p % TA;
            if (a0 + a >= n_across) continue;
            // the points past the stencil of the tile's last computed face feed no face: skip them
            // (a bounded line ends 3 ghost nodes past its last owned node; a periodic line shorter
            // than the tile would wrap more than once)
            if (s > nfa - b0 + 4) continue;
            const int along = along_shift(B, DIR, b0 - sh, s - 2, n_along);
            const uint32_t x = tile_node<DIR>(B, along, a0 + a, fixed);
            const S rho = ld(P0, x), u = ld(P1, x), v = ld(P2, x), w = ld(P3, x), pr = ld(P4, x);
            S q[5], f[5];
            prim_to_q(rho, u, v, w, pr, q);
            node_flux_e(rho, u, v, w, pr, q[4] + pr, ld(geo<S>(B, 1 + 3 * DIR), x), ld(geo<S>(B, 2 + 3 * DIR), x),
                        ld(geo<S>(B, 3 + 3 * DIR), x), ld(geo<S>(B, 0), x), f);
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                sq[lds_at<DIR, TA, TB>(c, s, a)] = q[c];
                sf[lds_at<DIR, TA, TB>(c, s, a)] = f[c];
            }
            if constexpr (STAGE_ROE) {
                sp[lds_at<DIR, TA, TB>(0, s, a)] = rho; sp[lds_at<DIR, TA, TB>(1, s, a)] = u;
                sp[lds_at<DIR, TA, TB>(2, s, a)] = v; sp[lds_at<DIR, TA, TB>(3, s, a)] = w;
                sp[lds_at<DIR, TA, TB>(4, s, a)] = enthalpy(rho, u, v, w, pr);
            }
        }
        __syncthreads();

        const int b = DIR == 0 ? tid % TB : tid / TA;
        const int a = DIR == 0 ? tid / TB : tid % TA;
        if (b0 + b < nfa && a0 + a < n_across) {
            S q[6][5];
#pragma unroll
            for (int m = 0; m < 6; ++m)
#pragma unroll
                for (int c = 0; c < 5; ++c) q[m][c] = sq[lds_at<DIR, TA, TB>(c, b + m, a)];
            const uint32_t gl = tile_node<DIR>(B, b0 + b - sh, a0 + a, fixed);
            S roe[5];
            if constexpr (STAGE_ROE) {
                // the face's two centre nodes are stencil points 2 and 3
                auto at = [&](int c, int m) { return sp[lds_at<DIR, TA, TB>(c, b + m, a)]; };
                roe_average(at(0, 2), at(1, 2), at(2, 2), at(3, 2), at(4, 2), at(0, 3), at(1, 3), at(2, 3), at(3, 3), at(4, 3), roe);
            } else {
                // from the conserved states of the two centre nodes, already in registers
                S ra, ua, va, wa, ha, rb, ub, vb, wb, hb;
                q_roe_inputs(q[2], ra, ua, va, wa, ha);
                q_roe_inputs(q[3], rb, ub, vb, wb, hb);
                roe_average(ra, ua, va, wa, ha, rb, ub, vb, wb, hb, roe);
            }
            // unit face normal and det_f |k| (the stored face arrays after geometry_finalize);
            // streaming accesses on periodic blocks (B.code null)
            S fg[4];
            if (B.code) {
#pragma unroll
                for (int c = 0; c < 3; ++c) fg[c] = ld(geo<S>(B, 10 + 3 * DIR + c), gl);
                fg[3] = ld(geo<S>(B, 19 + DIR), gl);
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) fg[c] = ldx<true>(geo<S>(B, 10 + 3 * DIR + c), gl);
                fg[3] = ldx<true>(geo<S>(B, 19 + DIR), gl);
            }
            S llf[5];
            face_llf(q, roe, fg[0], fg[1], fg[2], fg[3], S(B.eps), llf);
            S fh[5];
#pragma unroll
            for (int c = 0; c < 5; ++c)
                fh[c] = S(llf[c] + midpoint_correction(sf[lds_at<DIR, TA, TB>(c, b, a)], sf[lds_at<DIR, TA, TB>(c, b + 1, a)],
                                                       sf[lds_at<DIR, TA, TB>(c, b + 2, a)], sf[lds_at<DIR, TA, TB>(c, b + 3, a)],
                                                       sf[lds_at<DIR, TA, TB>(c, b + 4, a)], sf[lds_at<DIR, TA, TB>(c, b + 5, a)]));
            if (B.code) {
#pragma unroll
                for (int c = 0; c < 5; ++c) st((S *)B.FH[DIR][c], gl, fh[c]);
            } else {
#pragma unroll
                for (int c = 0; c < 5; ++c) stx<true>((S *)B.FH[DIR][c], gl, fh[c]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// update kernel: L and the SSP-RK3 stage (Weno.cl:3060-3135), in double.
//   mode 0..2: RK stage j; mode 3: write L to the block's Lout only
// ---------------------------------------------------------------------------
// L and the SSP-RK3 stage of one node; mdu / mdv / mde: the node's monitor values (mode 2 with
// the monitor on). LOC as in visc_node.
template <typename S, bool LOC, bool NT>
__device__ __forceinline__ void update_node(const BlockDev &B, int i, int j, int k, int mode, double &mdu, double &mdv,
                                            double &mde)
{
    const uint32_t g = node_index(B, i, j, k);
    double L[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const uint32_t m1 = shifted(B, i, j, k, d, -1);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const S *fh = (const S *)B.FH[d][c];
            L[c] += (double)ld(fh, g) - (double)ld(fh, m1);
        }
    }
    if (B.viscous) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const uint32_t m2 = shifted(B, i, j, k, d, -2), m1 = shifted(B, i, j, k, d, -1);
            const uint32_t p1 = shifted(B, i, j, k, d, 1), p2 = shifted(B, i, j, k, d, 2);
            if (LOC && near_side(B, i, j, k, d)) {
                // bounded block: the loc-selected forms of Weno.cl:2732-2856
                const uint32_t m3 = shifted(B, i, j, k, d, -3), p3 = shifted(B, i, j, k, d, 3);
                const int LI = ld(B.code, g), LM2 = ld(B.code, m2), LM1 = ld(B.code, m1), LP1 = ld(B.code, p1),
                          LP2 = ld(B.code, p2);
#pragma unroll
                for (int c = 1; c < 5; ++c) {
                    const S *fv = (const S *)B.Fv[d][c - 1];
                    const double f[7] = {(double)ld(fv, m3), (double)ld(fv, m2), (double)ld(fv, m1), (double)ld(fv, g),
                                         (double)ld(fv, p1), (double)ld(fv, p2), (double)ld(fv, p3)};
                    L[c] += loc_derivative(f, LI, LM2, LM1, LP1, LP2, true);
                }
            } else {
#pragma unroll
                for (int c = 1; c < 5; ++c) {
                    const S *f = (const S *)B.Fv[d][c - 1];
                    L[c] += (1.0 / 12.0) * (8.0 * ((double)ld(f, p1) - (double)ld(f, m1)) - ((double)ld(f, p2) - (double)ld(f, m2)));
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) L[c] = -L[c];
    if (mode == 3) {
#pragma unroll
        for (int c = 0; c < 5; ++c) B.Lout[(size_t)c * B.n + g] = L[c];
    } else {
        const double del_t = B.del_t;
        double un[5], u0[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) u0[c] = ldx<NT>(B.U0[c], g);
        if (mode == 0) {
#pragma unroll
            for (int c = 0; c < 5; ++c) { un[c] = u0[c] + del_t * L[c]; stx<NT>(B.Us[c], g, un[c]); }
        } else if (mode == 1) {
#pragma unroll
            for (int c = 0; c < 5; ++c) { un[c] = 0.75 * u0[c] + 0.25 * ldx<NT>(B.Us[c], g) + 0.25 * del_t * L[c]; stx<NT>(B.Us[c], g, un[c]); }
        } else {
#pragma unroll
            for (int c = 0; c < 5; ++c) un[c] = (u0[c] * (1.0 / 3.0)) + ((2.0 / 3.0) * ldx<NT>(B.Us[c], g)) + (2.0 / 3.0) * (del_t * L[c]);
        }
        const double det = ldx<NT>(B.det, g);
        const double idet = 1.0 / det;
        const double rho = idet * un[0], ir = 1.0 / rho;
        const double u = (idet * un[1]) * ir, v = (idet * un[2]) * ir, w = (idet * un[3]) * ir;
        const double e = (idet * un[4]) * ir - 0.5 * (u * u + v * v + w * w);
        const double p = e * rho * 0.4;
        stx<NT>((S *)B.P[0], g, S(rho)); stx<NT>((S *)B.P[1], g, S(u)); stx<NT>((S *)B.P[2], g, S(v)); stx<NT>((S *)B.P[3], g, S(w)); stx<NT>((S *)B.P[4], g, S(p));
        if (mode == 2) {
            // next step's RK base: U_Q recomputes det*Q from the new primitives (Solver.cpp:377-388)
            double q[5];
            prim_to_q(rho, u, v, w, p, q);
#pragma unroll
            for (int c = 0; c < 5; ++c) stx<NT>(B.U0[c], g, det * q[c]);
            if (B.mon) {
                const double r0 = idet * u0[0], ir0 = 1.0 / r0;
                const double uo = (idet * u0[1]) * ir0, vo = (idet * u0[2]) * ir0, wo = (idet * u0[3]) * ir0;
                const double eo = (idet * u0[4]) * ir0 - 0.5 * (uo * uo + vo * vo + wo * wo);
                mdu = fabs(u - uo); mdv = fabs(v - vo); mde = fabs(e - eo);
            }
        }
    }
}

// one maximum per wave of the monitor values (every lane takes part), written by lane 0 at slot
__device__ __forceinline__ void monitor_wave(const BlockDev &B, uint32_t slot, double mdu, double mdv, double mde)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mdu = fmax(mdu, __shfl_xor(mdu, o));
        mdv = fmax(mdv, __shfl_xor(mdv, o));
        mde = fmax(mde, __shfl_xor(mde, o));
    }
    if ((threadIdx.x & 63) == 0) {
        double *m = B.mon + 3 * (size_t)slot;
        m[0] = mdu; m[1] = mdv; m[2] = mde;
    }
}

// tiles as in k_visc; the shell tiles' monitor slots follow the block's node tiles' (tiles x
// waves per tile), one per wave of 64 shell nodes
template <typename S, int TX, int TY, bool BND>
__global__ void __launch_bounds__(TX *TY) k_update(const BlockDev *__restrict__ blks, const uint32_t *__restrict__ prefix,
                                                   uint32_t nblk, uint32_t total, int mode, uint32_t shell_tiles)
{
    const uint32_t all = total + (BND ? nblk * shell_tiles : 0u);
    for (uint32_t wg = blockIdx.x; wg < all; wg += gridDim.x) {
        // residual monitor of the step (Solver.cpp:546-573): |du|, |dv|, |de| against the RK base
        double mdu = 0.0, mdv = 0.0, mde = 0.0;
        if (BND && wg >= total) {
            const uint32_t s = wg - total;
            const BlockDev &B = blks[s / shell_tiles];
            const uint32_t t = (s % shell_tiles) * (TX * TY) + threadIdx.x;
            const uint32_t ns = shell_count(B);
            if ((s % shell_tiles) * (TX * TY) >= ns) continue;   // whole tile past the shell
            if (t < ns) {
                int i, j, k;
                shell_node(B, t, i, j, k);
                update_node<S, true, !BND>(B, i, j, k, mode, mdu, mdv, mde);
            }
            if (mode == 2 && B.mon && (t & ~63u) < ns) {
                const uint32_t main_slots = ((B.ni + TX - 1) / TX) * ((B.nj + TY - 1) / TY) * B.nk * (uint32_t)(TX * TY / 64);
                monitor_wave(B, main_slots + t / 64, mdu, mdv, mde);
            }
            continue;
        }
        const uint32_t t = xcd_tile(wg, total);
        const BlockDev &B = blks[find_block(prefix, nblk, t)];
        const uint32_t lt = t - B.tile0[0];
        int i0, j0, k;
        node_tile(B, lt, TX, TY, i0, j0, k);
        const int i = i0 + (int)(threadIdx.x % TX), j = j0 + (int)(threadIdx.x / TX);
        if (i < B.ni && j < B.nj && !(BND && B.code && in_shell(B, i, j, k)))
            update_node<S, false, !BND>(B, i, j, k, mode, mdu, mdv, mde);
        if (mode == 2 && B.mon) monitor_wave(B, lt * (TX * TY / 64) + threadIdx.x / 64, mdu, mdv, mde);
    }
}

// maxima over the per-wave monitor values: out[0..2] = max |du|, |dv|, |de|
__global__ void __launch_bounds__(256) k_mon_reduce(const double *__restrict__ mon, uint32_t count, double *__restrict__ out)
{
    __shared__ double sh[3][256];
    double a = 0.0, b = 0.0, c = 0.0;
    for (uint32_t q = threadIdx.x; q < count; q += 256) {
        a = fmax(a, mon[3 * q]); b = fmax(b, mon[3 * q + 1]); c = fmax(c, mon[3 * q + 2]);
    }
    sh[0][threadIdx.x] = a; sh[1][threadIdx.x] = b; sh[2][threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int r = 0; r < 3; ++r) sh[r][threadIdx.x] = fmax(sh[r][threadIdx.x], sh[r][threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x < 3) out[threadIdx.x] = sh[threadIdx.x][0];
}

// owned node o (i-fastest over ni x nj x nk) -> stored index
__device__ __forceinline__ uint32_t owned_node(const BlockDev &B, uint32_t o)
{
    const int i = (int)(o % (uint32_t)B.ni), j = (int)((o / (uint32_t)B.ni) % (uint32_t)B.nj), k = (int)(o / ((uint32_t)B.ni * B.nj));
    return node_index(B, i, j, k);
}

__device__ __forceinline__ uint32_t owned_count(const BlockDev &B) { return (uint32_t)B.ni * B.nj * B.nk; }

// set_flow: double primitives staged component-major in owned order (stage[c * n_owned + o]) ->
// P (type S) and U0 = det * Q(P) in double
template <typename S>
__global__ void k_base(BlockDev B, const double *__restrict__ stage)
{
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x, no = owned_count(B);
    if (o >= no) return;
    const uint32_t g = owned_node(B, o);
    const double rho = stage[o], u = stage[no + o], v = stage[2 * (size_t)no + o], w = stage[3 * (size_t)no + o],
                 p = stage[4 * (size_t)no + o];
    double q[5];
    prim_to_q(rho, u, v, w, p, q);
    const double det = ld(B.det, g);
#pragma unroll
    for (int c = 0; c < 5; ++c) st(B.U0[c], g, det * q[c]);
    st((S *)B.P[0], g, S(rho)); st((S *)B.P[1], g, S(u)); st((S *)B.P[2], g, S(v)); st((S *)B.P[3], g, S(w)); st((S *)B.P[4], g, S(p));
}

// get_flow of the mixed form: primitives from the double RK base U0 / det, owned order
__global__ void k_prims(BlockDev B, double *__restrict__ out)
{
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x, no = owned_count(B);
    if (o >= no) return;
    const uint32_t g = owned_node(B, o);
    const double idet = 1.0 / ld(B.det, g);
    const double rho = idet * ld(B.U0[0], g), ir = 1.0 / rho;
    const double u = (idet * ld(B.U0[1], g)) * ir, v = (idet * ld(B.U0[2], g)) * ir, w = (idet * ld(B.U0[3], g)) * ir;
    const double e = (idet * ld(B.U0[4], g)) * ir - 0.5 * (u * u + v * v + w * w);
    out[o] = rho; out[no + o] = u; out[2 * (size_t)no + o] = v; out[3 * (size_t)no + o] = w; out[4 * (size_t)no + o] = e * rho * 0.4;
}

// get_flow of the plain form: the primitives of the last stage, owned order
__global__ void k_copy_prims(BlockDev B, double *__restrict__ out)
{
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x, no = owned_count(B);
    if (o >= no) return;
    const uint32_t g = owned_node(B, o);
#pragma unroll
    for (int c = 0; c < 5; ++c) out[(size_t)c * no + o] = ld((const double *)B.P[c], g);
}

// ---------------------------------------------------------------------------
// side conditions of bounded blocks (Flow_Initialisation.cpp face rules; the C-grid wake cut)
// ---------------------------------------------------------------------------
template <typename S>
__device__ __forceinline__ void get_prim(const BlockDev &B, uint32_t g, double (&w)[5])
{
#pragma unroll
    for (int c = 0; c < 5; ++c) w[c] = (double)ld((const S *)B.P[c], g);
}

// primitives of a boundary or ghost node; its RK base (mode 2) or stage vector (modes 0, 1)
// follows from them, as U_Q rebuilds U from FLOW (ghost stage vectors are not defined by the
// reference at stages 1 and 2; here they are det * Q of the side condition's state). The
// primitives are rounded to S first, as if read back from the stored arrays.
template <typename S>
__device__ __forceinline__ void put_node(const BlockDev &B, uint32_t g, const double (&w)[5], int mode, bool ghost)
{
    double r[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) r[c] = (double)S(w[c]);
    const bool u = mode == 2 || ghost;
    const double det = u ? ld(B.det, g) : 0.0;
#pragma unroll
    for (int c = 0; c < 5; ++c) st((S *)B.P[c], g, S(w[c]));
    if (!u) return;
    double q[5];
    prim_to_q(r[0], r[1], r[2], r[3], r[4], q);
#pragma unroll
    for (int c = 0; c < 5; ++c) st(mode == 2 ? B.U0[c] : B.Us[c], g, det * q[c]);
}

// the node across a wake cut into a ghost or duplicate node: primitives copied, the stage vector
// and RK base carry the same Q with the destination's own det (U = det Q). Every value is loaded
// before the first store.
template <typename S>
__device__ __forceinline__ void copy_node(const BlockDev &B, uint32_t dst, uint32_t src)
{
    const double ratio = ld(B.det, dst) / ld(B.det, src);
    S p[5];
    double us[5], u0[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        p[c] = ld((const S *)B.P[c], src);
        us[c] = ld(B.Us[c], src);
        u0[c] = ld(B.U0[c], src);
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        st((S *)B.P[c], dst, p[c]);
        st(B.Us[c], dst, us[c] * ratio);
        st(B.U0[c], dst, u0[c] * ratio);
    }
}

// The two sides of direction d (blockIdx.z = 0: low side, 1: high side), one thread per node of
// a side's face and layer m (0: the boundary node, 1..3: the ghost layers). The nodes a side
// writes are its boundary nodes and ghost layers; it reads the first owned layers in from its
// side (a bounded direction has >= 7 nodes), so the two sides of a direction are independent and
// the directions run one after the other (i, then j, then k: a node on two sides ends with the
// later direction's value).
template <typename S>
__global__ void __launch_bounds__(256) k_bc(const BlockDev *__restrict__ blks, int d, int mode)
{
    const BlockDev &B = blks[blockIdx.y];
    if (!B.off[d]) return;
    const int hi = (int)blockIdx.z, side = 2 * d + hi;
    const int na = d == 0 ? B.nj : B.ni, nb = d == 2 ? B.nj : B.nk;
    const uint32_t nf = (uint32_t)na * (uint32_t)nb;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 4 * nf) return;
    // i sides: the four layers of a face node are adjacent in memory, so they go to four adjacent
    // threads; j and k sides: adjacent threads take adjacent i of one layer
    const int m = d == 0 ? (int)(t & 3u) : (int)(t / nf);
    const uint32_t r = d == 0 ? t >> 2 : t - (uint32_t)m * nf;
    const int a = (int)(r % (uint32_t)na), bq = (int)(r / (uint32_t)na);
    const int n = d == 0 ? B.ni : d == 1 ? B.nj : B.nk;
    auto at = [&](int c) -> uint32_t {
        if (d == 0) return node_index(B, c, a, bq);
        if (d == 1) return node_index(B, a, c, bq);
        return node_index(B, a, bq, c);
    };
    const int c0 = hi ? n - 1 : 0, in = hi ? -1 : 1;   // boundary node; step into the block
    int kind = B.bc[side];
    if (kind == CFDX_BC_WALL_CUT) {
        if (a >= B.cut_i0 && a <= B.cut_i1) {
            kind = CFDX_BC_WALL;
        } else {
            // wake cut (j- side, a = i): node (i, -m) is node (ni-1-i, m); row 0 of the lower half
            // copies the upper half
            const int im = B.ni - 1 - a;
            if (m > 0) copy_node<S>(B, node_index(B, a, -m, bq), node_index(B, im, m, bq));
            else if (a < im) copy_node<S>(B, node_index(B, a, 0, bq), node_index(B, im, 0, bq));
            return;
        }
    }
    const double M2 = 1.4 * B.mach * B.mach;
    double w[5];
    if (kind == CFDX_BC_WALL || kind == CFDX_BC_EXTRAP) {
        get_prim<S>(B, at(c0 + in * (m ? m : 1)), w);
        if (m == 0) {
            if (kind == CFDX_BC_WALL) {
                const double t1 = M2 * (w[4] / w[0]);
                w[0] = M2 * (w[4] / t1); w[1] = 0.0; w[2] = 0.0; w[3] = 0.0;
            }
        } else if (kind == CFDX_BC_WALL) {
            w[1] = -w[1]; w[2] = -w[2]; w[3] = -w[3];
        }
    } else if (kind == CFDX_BC_FARFIELD) {
        const double p = 1.0 / M2;
        w[0] = M2 * (p / 1.0); w[1] = B.cos_aoa; w[2] = B.sin_aoa; w[3] = 0.0; w[4] = p;
    } else {   // CFDX_BC_OUTFLOW
        get_prim<S>(B, at(c0 + in), w);
        const double t1 = M2 * (w[4] / w[0]);
        w[0] = M2 * (w[4] / t1);
    }
    put_node<S>(B, at(c0 - in * m), w, mode, m > 0);
}

// viscous node fluxes of the ghost rows across a wake cut: the mirror node's, with the signs of
// the i and j directions reversed (zeta and eta run the other way on the other side of the cut);
// one thread per (i, layer m, k)
template <typename S>
__global__ void __launch_bounds__(256) k_cut_viscous(const BlockDev *__restrict__ blks)
{
    const BlockDev &B = blks[blockIdx.y];
    if (!B.off[1] || B.bc[2] != CFDX_BC_WALL_CUT || !B.viscous) return;
    const uint32_t nf = (uint32_t)B.ni * (uint32_t)B.nk;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nf) return;
    const int m = 1 + (int)(t / nf);
    const uint32_t r = t % nf;
    const int i = (int)(r % (uint32_t)B.ni), k = (int)(r / (uint32_t)B.ni);
    if (i >= B.cut_i0 && i <= B.cut_i1) return;
    const uint32_t dst = node_index(B, i, -m, k), src = node_index(B, B.ni - 1 - i, m, k);
    S v[3][4];
#pragma unroll
    for (int dd = 0; dd < 3; ++dd)
#pragma unroll
        for (int c = 0; c < 4; ++c) v[dd][c] = ld((const S *)B.Fv[dd][c], src);
#pragma unroll
    for (int dd = 0; dd < 3; ++dd)
#pragma unroll
        for (int c = 0; c < 4; ++c) st((S *)B.Fv[dd][c], dst, dd == 2 ? v[dd][c] : -v[dd][c]);
}

template <typename S>
void launch_bc(const LaunchList &L, int mode, hipStream_t st)
{
    for (int d = 0; d < 3; ++d)
        if (L.bc_face_d[d])
            hipLaunchKernelGGL(k_bc<S>, dim3((4 * L.bc_face_d[d] + 255) / 256, L.nblk, 2), dim3(256), 0, st, L.blks, d, mode);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
template <typename S, int V>
void launch_face_i(const LaunchList &L, dim3 g, hipStream_t st)
{
    hipLaunchKernelGGL((k_face<S, 0, FACE_I_TILES[V][0], FACE_I_TILES[V][1]>), g, dim3(256), 0, st, L.blks, L.prefix[1],
                       L.nblk, L.total[1]);
}
template <typename S, int DIR, int V>
void launch_face_jk(const LaunchList &L, dim3 g, hipStream_t st)
{
    hipLaunchKernelGGL((k_face<S, DIR, FACE_JK_TILES[V][0], FACE_JK_TILES[V][1]>), g, dim3(256), 0, st, L.blks,
                       L.prefix[1 + DIR], L.nblk, L.total[1 + DIR]);
}
template <typename S, int V, bool BND>
void launch_node_b(const LaunchList &L, dim3 g, hipStream_t st, bool visc_kernel, int mode)
{
    constexpr int TX = NODE_TILES[V][0], TY = NODE_TILES[V][1];
    static_assert(TX * TY == 256, "node tiles have 256 threads");
    const uint32_t shell_tiles = BND ? (L.shell_max + TX * TY - 1) / (TX * TY) : 0u;
    dim3 ga = g;
    if (BND && !L.grid_cap) ga.x += L.nblk * shell_tiles;   // one workgroup per shell tile too
    if (visc_kernel)
        hipLaunchKernelGGL((k_visc<S, TX, TY, BND>), ga, dim3(TX * TY), 0, st, L.blks, L.prefix[0], L.nblk, L.total[0], mode,
                           shell_tiles);
    else
        hipLaunchKernelGGL((k_update<S, TX, TY, BND>), ga, dim3(TX * TY), 0, st, L.blks, L.prefix[0], L.nblk, L.total[0], mode,
                           shell_tiles);
}
// lists without bounded blocks run the node kernels without the loc-code branches
template <typename S, int V>
void launch_node(const LaunchList &L, dim3 g, hipStream_t st, bool visc_kernel, int mode)
{
    if (L.bc_face) launch_node_b<S, V, true>(L, g, st, visc_kernel, mode);
    else launch_node_b<S, V, false>(L, g, st, visc_kernel, mode);
}

template <typename S>
void node_kernel(const LaunchList &L, dim3 g, hipStream_t st, bool visc_kernel, int mode)
{
    switch (L.tiles.node) {
    case 1: launch_node<S, 1>(L, g, st, visc_kernel, mode); break;
    case 2: launch_node<S, 2>(L, g, st, visc_kernel, mode); break;
    case 3: launch_node<S, 3>(L, g, st, visc_kernel, mode); break;
    default: launch_node<S, 0>(L, g, st, visc_kernel, mode);
    }
}

template <typename S, int DIR>
void face_jk(const LaunchList &L, dim3 g, hipStream_t st)
{
    switch (L.tiles.face[DIR]) {
    case 1: launch_face_jk<S, DIR, 1>(L, g, st); break;
    case 2: launch_face_jk<S, DIR, 2>(L, g, st); break;
    case 3: launch_face_jk<S, DIR, 3>(L, g, st); break;
    default: launch_face_jk<S, DIR, 0>(L, g, st);
    }
}

template <typename S>
void launch_stage(const LaunchList &L, int mode, hipStream_t st, bool any_visc, hipEvent_t *ev, const Overlap *ov)
{
    auto rec = [&](int s) { if (ev) (void)hipEventRecord(ev[s], st); };
    auto grid = [&](uint32_t total) { return dim3(L.grid_cap ? std::min(total, L.grid_cap) : total); };
    rec(0);
    // the viscous node kernel reads the stage vector, primitives and geometry and writes Fv and mu;
    // the face sweeps read the primitives and geometry and write F-hat: independent until the
    // update, so the viscous kernel may run beside the sweeps on the second stream (persistent,
    // one workgroup per CU, leaving the CUs' remaining slots to the sweeps)
    const bool overlap = ov && !ev && any_visc;
    hipStream_t vs = st;
    dim3 gv = grid(L.total[0]);
    if (overlap) {
        (void)hipEventRecord(ov->fork, st);
        (void)hipStreamWaitEvent(ov->aux, ov->fork, 0);
        vs = ov->aux;
        gv.x = std::min(gv.x, ov->visc_grid);
    }
    if (any_visc) node_kernel<S>(L, gv, vs, true, mode);
    if (any_visc && L.any_cut)
        hipLaunchKernelGGL(k_cut_viscous<S>, dim3((3 * L.bc_face_d[1] + 255) / 256, L.nblk), dim3(256), 0, vs, L.blks);
    if (overlap) (void)hipEventRecord(ov->join, vs);
    rec(1);
    switch (L.tiles.face[0]) {
    case 1: launch_face_i<S, 1>(L, grid(L.total[1]), st); break;
    case 2: launch_face_i<S, 2>(L, grid(L.total[1]), st); break;
    default: launch_face_i<S, 0>(L, grid(L.total[1]), st);
    }
    rec(2);
    face_jk<S, 1>(L, grid(L.total[2]), st);
    rec(3);
    face_jk<S, 2>(L, grid(L.total[3]), st);
    rec(4);
    if (overlap) (void)hipStreamWaitEvent(st, ov->join, 0);
    node_kernel<S>(L, grid(L.total[0]), st, false, mode);
    if (L.bc_face && mode <= 2) launch_bc<S>(L, mode, st);
    rec(5);
}

void stage(const LaunchList &L, int precision, int mode, hipStream_t st, bool any_visc, hipEvent_t *ev, const Overlap *ov)
{
    if (precision == CFDX_MIXED) launch_stage<float>(L, mode, st, any_visc, ev, ov);
    else launch_stage<double>(L, mode, st, any_visc, ev, ov);
}

void base(const BlockDev &B, int precision, hipStream_t st, const double *stage)
{
    const unsigned nb = (unsigned)(((uint32_t)B.ni * B.nj * B.nk + 255) / 256);
    if (precision == CFDX_MIXED) hipLaunchKernelGGL(k_base<float>, dim3(nb), dim3(256), 0, st, B, stage);
    else hipLaunchKernelGGL(k_base<double>, dim3(nb), dim3(256), 0, st, B, stage);
}

void boundary(const LaunchList &L, int precision, int mode, hipStream_t st)
{
    if (!L.bc_face) return;
    if (precision == CFDX_MIXED) launch_bc<float>(L, mode, st);
    else launch_bc<double>(L, mode, st);
}

// get_flow: the reference's FLOW record of every owned node, owned order (Weno.cl:3126-3135 for
// t, a, e; mu of the last viscous pass). The primitives are those of the last stage (plain form)
// or recovered in double from the RK base (mixed form).
__global__ void k_flow_out(BlockDev B, cfdx_flow *__restrict__ out, double mach, int from_base)
{
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x, no = owned_count(B);
    if (o >= no) return;
    const uint32_t g = owned_node(B, o);
    double rho, u, v, w, p;
    if (from_base) {
        const double idet = 1.0 / ld(B.det, g);
        rho = idet * ld(B.U0[0], g);
        const double ir = 1.0 / rho;
        u = (idet * ld(B.U0[1], g)) * ir; v = (idet * ld(B.U0[2], g)) * ir; w = (idet * ld(B.U0[3], g)) * ir;
        const double e = (idet * ld(B.U0[4], g)) * ir - 0.5 * (u * u + v * v + w * w);
        p = e * rho * 0.4;
    } else {
        rho = ld((const double *)B.P[0], g); u = ld((const double *)B.P[1], g); v = ld((const double *)B.P[2], g);
        w = ld((const double *)B.P[3], g); p = ld((const double *)B.P[4], g);
    }
    // the record's 9 doubles in order u, v, w, p, t, rho, mu, a, e (out is 8-byte aligned)
    double *r = (double *)out + 9 * (size_t)o;
    r[0] = u; r[1] = v; r[2] = w; r[3] = p;
    r[4] = (1.4 * mach * mach) * (p / rho);
    r[5] = rho;
    r[6] = ld(B.mu, g);
    r[7] = sqrt(1.4 * p / rho);
    r[8] = p / (0.4 * rho);
}

void flow_out(const BlockDev &B, int precision, hipStream_t st, cfdx_flow *out, double mach)
{
    const unsigned nb = (unsigned)(((uint32_t)B.ni * B.nj * B.nk + 255) / 256);
    hipLaunchKernelGGL(k_flow_out, dim3(nb), dim3(256), 0, st, B, out, mach, precision == CFDX_MIXED ? 1 : 0);
}

// number of workgroups that met a value that is not finite among p[0..n)
__global__ void __launch_bounds__(256) k_nonfinite(const double *__restrict__ p, size_t n, unsigned *__restrict__ count)
{
    bool bad = false;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        bad |= !__builtin_isfinite(p[q]);
    if (__syncthreads_or(bad) && threadIdx.x == 0) atomicAdd(count, 1u);
}

void count_nonfinite(const double *p, size_t n, unsigned *count, hipStream_t st)
{
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_nonfinite, dim3(nb), dim3(256), 0, st, p, n, count);
}

__global__ void __launch_bounds__(256) k_geo32(const double *__restrict__ src, float *__restrict__ dst, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        dst[q] = (float)src[q];
}

// The face arrays (metric row k of the +1/2 face along each direction, arrays 10-18, and its det
// det_f, arrays 19-21) into the form the face sweeps use: the unit normal k / |k| and det_f |k|.
// A stored node with no face (k = 0) keeps zeros.
__global__ void __launch_bounds__(256) k_unit_normals(double *__restrict__ geo, size_t n)
{
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double *k = geo + (size_t)(10 + 3 * d) * n + g, *df = geo + (size_t)(19 + d) * n + g;
            const double a = k[0], b = k[n], c = k[2 * n];
            const double kn = sqrt(a * a + b * b + c * c);
            if (kn > 0.0) {
                k[0] = a / kn;
                k[n] = b / kn;
                k[2 * n] = c / kn;
                df[0] = df[0] * kn;
            }
        }
    }
}

void geometry_finalize(const BlockDev &B, hipStream_t st)
{
    const size_t n = B.n;
    const unsigned nf = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_unit_normals, dim3(nf), dim3(256), 0, st, const_cast<double *>(B.det), n);
    if (!B.geo32) return;
    const size_t n22 = (size_t)22 * B.n;
    const unsigned nb = (unsigned)std::min<size_t>((n22 + 255) / 256, 8192);
    hipLaunchKernelGGL(k_geo32, dim3(nb), dim3(256), 0, st, B.det, const_cast<float *>(B.geo32), n22);
}

void prims(const BlockDev &B, int precision, hipStream_t st, double *out)
{
    const unsigned nb = (unsigned)(((uint32_t)B.ni * B.nj * B.nk + 255) / 256);
    if (precision == CFDX_MIXED) hipLaunchKernelGGL(k_prims, dim3(nb), dim3(256), 0, st, B, out);
    else hipLaunchKernelGGL(k_copy_prims, dim3(nb), dim3(256), 0, st, B, out);
}

void monitor_reduce(const double *mon, uint32_t count, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_mon_reduce, dim3(1), dim3(256), 0, st, mon, count, out);
}

uint32_t tiles_of(const BlockDev &B, int kind, const TileCfg &cfg)
{
    if (kind == 0) {
        const int TX = NODE_TILES[cfg.node][0], TY = NODE_TILES[cfg.node][1];
        return (uint32_t)(((B.ni + TX - 1) / TX) * ((B.nj + TY - 1) / TY) * B.nk);
    }
    auto f = [&](const int (&t)[2], int along, int across, int fix) {
        return (uint32_t)(((across + t[0] - 1) / t[0]) * ((along + t[1] - 1) / t[1]) * fix);
    };
    if (kind == 1) return f(FACE_I_TILES[cfg.face[0]], faces_along(B, 0), B.nj, B.nk);
    if (kind == 2) return f(FACE_JK_TILES[cfg.face[1]], faces_along(B, 1), B.ni, B.nk);
    return f(FACE_JK_TILES[cfg.face[2]], faces_along(B, 2), B.ni, B.nj);
}

}  // namespace blk
}  // namespace cfdx
