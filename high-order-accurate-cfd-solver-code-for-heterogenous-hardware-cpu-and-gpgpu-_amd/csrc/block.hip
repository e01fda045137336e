// Host objects of the structured fast path: cfdx_block (one periodic grid resident in HBM)
// and cfdx_batch (a list of blocks stepped together, one launch per kernel class for all of
// them, optionally replayed from a captured hipGraph of one RK3 step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "block.hpp"

using cfdx::blk::BlockDev;
using cfdx::blk::LaunchList;

struct cfdx_block {
    cfdx_ctx *ctx = nullptr;
    cfdx_block_desc desc{};
    BlockDev B{};                 // host image of the device descriptor
    void *mem = nullptr;          // one allocation for every array of the block
    size_t bytes = 0;             // the arrays (bytes_of)
    size_t alloc = 0;             // the allocation: arrays, then dB and dprefix
    BlockDev *dB = nullptr;       // device copy of B (single-block launches)
    uint32_t *dprefix = nullptr;  // 4 x 2 tile offsets of the single-block launch list
    uint64_t version = 1;         // bumped whenever B changes
    uint64_t uploaded = 0;        // version of B in dB
    bool configured = false;
    bool check = false;           // non-finite check of the state after each cfdx_block_step
    long long steps_done = 0;     // steps taken by cfdx_block_step (for the check's message)
    double last_mach = 0.5;
    bool timing = false;
    hipEvent_t ev[6] = {};
    double t_visc = 0, t_face[3] = {0, 0, 0}, t_update = 0;
    int launches = 0;
    cfdx::blk::TileCfg tiles;
    double *mon = nullptr;        // residual monitor partials (3 doubles per wave) + 3 reduced values
    uint32_t mon_cap = 0;         // partial triples allocated
    uint32_t mon_count = 0;       // partial triples written by the last step
    std::vector<cfdx_batch *> batches;   // batches holding the block (it leaves them when destroyed)
    // recorded on a batch's stream when the block leaves it (cfdx_batch_remove): later work on the
    // block waits for the batch's last launches over it, which may run on another context's stream
    // (left_pending); cfdx_batch_add records it on the block's stream for the batch to wait on
    hipEvent_t left = nullptr;
    bool left_pending = false;
};

struct cfdx_batch {
    cfdx_ctx *ctx = nullptr;
    int precision = CFDX_FP64;
    std::vector<cfdx_block *> blocks;
    std::vector<uint64_t> seen;   // block versions in dB
    BlockDev *dB = nullptr;
    uint32_t *dprefix = nullptr;
    size_t cap = 0;
    bool dirty = true;
    LaunchList L;
    bool any_visc = false;
    uint32_t grid_cap = 0;
    cfdx::blk::TileCfg tiles;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t graph_gen = 0;       // context generation the graph was captured under
};

namespace {

// Orders work on a block's own stream after the launches of a batch it has left.
int wait_batches(cfdx_block *b)
{
    if (!b->left_pending) return CFDX_OK;
    // launches that have completed need no stream dependency (and a dependency on the batch's
    // stream while another thread captures it is refused)
    const hipError_t q = hipEventQuery(b->left);
    if (q == hipSuccess) {
        b->left_pending = false;
        return CFDX_OK;
    }
    if (q != hipErrorNotReady) return cfdx::hip_fail(q, "query the batch the block left");
    CFDX_HIP(hipStreamWaitEvent(b->ctx->stream, b->left, 0), "wait for the batch the block left");
    return CFDX_OK;
}

// enter the block's context; its stream waits for a batch the block has left
int enter_block(cfdx_block *b)
{
    CFDX_TRY(cfdx::enter(b->ctx));
    return wait_batches(b);
}

size_t flux_size(int precision) { return precision == CFDX_MIXED ? sizeof(float) : sizeof(double); }

// padded sizes and owned offsets: 3 ghost layers on each side of a bounded direction
void padded_dims(const cfdx_block_desc *d, int dims[3], int off[3])
{
    const int n[3] = {d->ni, d->nj, d->nk};
    for (int q = 0; q < 3; ++q) {
        off[q] = ((d->periodic >> q) & 1) ? 0 : 3;
        dims[q] = n[q] + 2 * off[q];
    }
}

int64_t stored_nodes(const cfdx_block_desc *d)
{
    int dims[3], off[3];
    padded_dims(d, dims, off);
    return (int64_t)dims[0] * dims[1] * dims[2];
}

// the block descriptor and the single-block prefix table after the arrays
size_t tail_bytes() { return ((sizeof(BlockDev) + 255) / 256) * 256 + 256; }

int64_t bytes_of(const cfdx_block_desc *d)
{
    const int64_t n = stored_nodes(d);
    // 33 double arrays (geometry 22, U0 5, Us 5, mu 1) + 32 flux-type arrays (P 5, Fv 12, FH 15)
    // + in a mixed block float copies of the 22 geometry arrays + one byte of loc code per node
    // for a bounded block
    const int64_t codes = d->periodic == 7 ? 0 : ((n + 255) / 256) * 256;
    const int64_t geo32 = d->precision == CFDX_MIXED ? 22 * 4 : 0;
    return n * (33 * 8 + 32 * (int64_t)flux_size(d->precision) + geo32) + codes + 256;
}

int check_desc(const cfdx_block_desc *d)
{
    if (d->ni < 3 || d->nj < 3 || d->nk < 3) return cfdx::fail(CFDX_E_INVALID, "block_create: each dimension must be >= 3");
    if (d->periodic < 0 || d->periodic > 7) return cfdx::fail(CFDX_E_INVALID, "block_create: periodic is a 3-bit mask, got %d", d->periodic);
    if (d->precision != CFDX_FP64 && d->precision != CFDX_MIXED)
        return cfdx::fail(CFDX_E_INVALID, "block_create: precision %d is neither CFDX_FP64 nor CFDX_MIXED", d->precision);
    const int n[3] = {d->ni, d->nj, d->nk};
    for (int q = 0; q < 3; ++q) {
        if ((d->periodic >> q) & 1) continue;
        if (n[q] < 7) return cfdx::fail(CFDX_E_INVALID, "block_create: a bounded direction needs >= 7 nodes (direction %d has %d)", q, n[q]);
        for (int h = 0; h < 2; ++h) {
            const int bc = d->bc[2 * q + h];
            if (bc < CFDX_BC_WALL || bc > CFDX_BC_WALL_CUT)
                return cfdx::fail(CFDX_E_INVALID, "block_create: side %d of a bounded direction has no condition (%d)", 2 * q + h, bc);
            if (bc == CFDX_BC_WALL_CUT && 2 * q + h != 2)
                return cfdx::fail(CFDX_E_INVALID, "block_create: CFDX_BC_WALL_CUT is a j- side condition");
        }
    }
    if (!(d->periodic & 2) && d->bc[2] == CFDX_BC_WALL_CUT &&
        !(d->cut_i0 >= 0 && d->cut_i0 <= d->cut_i1 && d->cut_i1 < d->ni && d->cut_i0 + d->cut_i1 == d->ni - 1))
        return cfdx::fail(CFDX_E_INVALID, "block_create: the wall range of a wake cut must be symmetric in i (cut_i0 + cut_i1 = ni - 1)");
    const long long np = stored_nodes(d);
    if (np >= (1LL << 29)) return cfdx::fail(CFDX_E_CAPACITY, "block_create: %lld stored nodes, the limit is 2^29 (32-bit byte offsets)", np);
    return CFDX_OK;
}

// loc codes of a bounded block: 0 interior, 1 node on a side (boundary condition node), 100 ghost;
// a ghost across a wake cut stands for its mirror node and takes the mirror's code
std::vector<uint8_t> loc_codes(const cfdx_block_desc *d)
{
    int dims[3], off[3];
    padded_dims(d, dims, off);
    const int n[3] = {d->ni, d->nj, d->nk};
    const bool cut = !(d->periodic & 2) && d->bc[2] == CFDX_BC_WALL_CUT;
    auto wall_col = [&](int i) { return i >= d->cut_i0 && i <= d->cut_i1; };
    std::vector<uint8_t> c((size_t)dims[0] * dims[1] * dims[2], 100);
    auto at = [&](int i, int j, int k) {
        return (size_t)(i + off[0]) + (size_t)dims[0] * ((size_t)(j + off[1]) + (size_t)dims[1] * (size_t)(k + off[2]));
    };
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const int x[3] = {i, j, k};
                bool side = false;
                for (int q = 0; q < 3; ++q) {
                    if (!off[q]) continue;
                    if (x[q] == n[q] - 1) side = true;
                    if (x[q] == 0 && !(q == 1 && cut && !wall_col(i))) side = true;
                }
                c[at(i, j, k)] = side ? 1 : 0;
            }
    if (cut)
        for (int k = 0; k < n[2]; ++k)
            for (int i = 0; i < n[0]; ++i)
                if (!wall_col(i))
                    for (int m = 1; m <= 3; ++m) c[at(i, -m, k)] = c[at(n[0] - 1 - i, m, k)];
    return c;
}

int check_params(const cfdx_params *p)
{
    if (!p) return cfdx::fail(CFDX_E_INVALID, "null params");
    if (p->sgs != CFDX_SGS_MTS && p->sgs != CFDX_SGS_SMAGORINSKY)
        return cfdx::fail(CFDX_E_INVALID, "sgs %d is neither CFDX_SGS_MTS nor CFDX_SGS_SMAGORINSKY", p->sgs);
    if (!(p->c_s >= 0)) return cfdx::fail(CFDX_E_INVALID, "c_s must be >= 0");
    const bool sgs_off = p->sgs == CFDX_SGS_MTS ? p->c_mts == 0.0 : p->c_s == 0.0;
    if (!p->viscous && !(std::isinf(p->reyl) && sgs_off))
        return cfdx::fail(CFDX_E_INVALID, "viscous=0 requires reyl=inf and the SGS model off (else the RHS differs)");
    if (!(p->eps > 0)) return cfdx::fail(CFDX_E_INVALID, "eps must be > 0");
    return CFDX_OK;
}

void configure(cfdx_block *b, double mach, double del_t, const cfdx_params *p)
{
    const BlockDev old = b->B;
    b->B.del_t = del_t;
    b->B.mach = mach;
    b->B.eps = p->eps;
    b->B.viscous = p->viscous ? 1 : 0;
    b->B.vp = cfdx::dev::make_visc_params(p->mach, p->reyl, p->pr, p->free_t, p->c_mts, p->c_t, p->sgs, p->c_s);
    b->last_mach = mach;
    b->configured = true;
    if (std::memcmp(&old, &b->B, sizeof old) != 0) ++b->version;
}

// values of the owned nodes (owned order) from an array over the stored nodes
void gather_owned(const cfdx_block *b, const double *stored, double *owned)
{
    const int ni = b->desc.ni, nj = b->desc.nj, nk = b->desc.nk;
    const int *off = b->B.off;
    size_t o = 0;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i)
                owned[o++] = stored[(size_t)(i + off[0]) + (size_t)b->B.pi * ((size_t)(j + off[1]) + (size_t)b->B.pj * (size_t)(k + off[2]))];
}

// side-condition extent of one block (largest side face of a bounded direction) and wake cut
void bc_extent(const BlockDev &B, LaunchList &L)
{
    if (!B.code) return;
    const uint32_t f[3] = {(uint32_t)B.nj * B.nk, (uint32_t)B.ni * B.nk, (uint32_t)B.ni * B.nj};
    for (int d = 0; d < 3; ++d)
        if (B.off[d]) {
            L.bc_face_d[d] = std::max(L.bc_face_d[d], f[d]);
            L.bc_face = std::max(L.bc_face, f[d]);
        }
    if (B.off[1] && B.bc[2] == CFDX_BC_WALL_CUT) L.any_cut = true;
    L.shell_max = std::max(L.shell_max, cfdx::blk::shell_count(B));
}

// single-block launch list (tile offsets 0 / total)
int upload_single(cfdx_block *b)
{
    if (b->uploaded == b->version) return CFDX_OK;
    BlockDev B = b->B;
    B.strip = (uint32_t)b->tiles.strip;
    uint32_t pre[8];
    for (int kind = 0; kind < 4; ++kind) {
        B.tile0[kind] = 0;
        pre[2 * kind] = 0;
        pre[2 * kind + 1] = cfdx::blk::tiles_of(b->B, kind, b->tiles);
    }
    CFDX_HIP(hipMemcpyAsync(b->dB, &B, sizeof B, hipMemcpyHostToDevice, b->ctx->stream), "block descriptor upload");
    CFDX_HIP(hipMemcpyAsync(b->dprefix, pre, sizeof pre, hipMemcpyHostToDevice, b->ctx->stream), "block prefix upload");
    CFDX_HIP(hipStreamSynchronize(b->ctx->stream), "block descriptor sync");
    b->uploaded = b->version;
    return CFDX_OK;
}

LaunchList single_list(cfdx_block *b)
{
    LaunchList L;
    L.blks = b->dB;
    L.nblk = 1;
    L.tiles = b->tiles;
    bc_extent(b->B, L);
    for (int kind = 0; kind < 4; ++kind) {
        L.prefix[kind] = b->dprefix + 2 * kind;
        L.total[kind] = cfdx::blk::tiles_of(b->B, kind, b->tiles);
    }
    return L;
}

uint32_t mon_triples(const BlockDev &B, const cfdx::blk::TileCfg &t)
{
    const int tx = cfdx::blk::NODE_TILES[t.node][0], ty = cfdx::blk::NODE_TILES[t.node][1];
    // node kernel waves, then one slot per wave of the shell kernel
    return cfdx::blk::tiles_of(B, 0, t) * (uint32_t)(tx * ty / 64) + (cfdx::blk::shell_count(B) + 63) / 64;
}

// the overlap of a stage's viscous kernel with its face sweeps on a context that allows it: for
// mixed-precision work always; for FP64 work only inside a captured graph (the FP64 sweeps hold
// 3 waves per SIMD at ~160 VGPRs, which leaves few registers for a viscous wave beside them:
// with stream events the overlap measured no gain, as graph edges 1 % at 128^3)
cfdx::blk::Overlap overlap_of(const cfdx_ctx *ctx)
{
    return cfdx::blk::Overlap{ctx->aux, ctx->fork, ctx->join, (uint32_t)std::max(ctx->n_cu, 1)};
}

const cfdx::blk::Overlap *overlap_if(const cfdx_ctx *ctx, int precision, const cfdx::blk::Overlap &ov,
                                     bool graph = false)
{
    return (ctx->overlap && (precision == CFDX_MIXED || graph)) ? &ov : nullptr;
}

int run_stage(cfdx_block *b, int mode)
{
    const LaunchList L = single_list(b);
    if (mode == 2) b->mon_count = mon_triples(b->B, b->tiles);
    const cfdx::blk::Overlap ov = overlap_of(b->ctx);
    cfdx::blk::stage(L, b->desc.precision, mode, b->ctx->stream, b->B.viscous != 0, b->timing ? b->ev : nullptr,
                     overlap_if(b->ctx, b->desc.precision, ov));
    b->launches += (b->B.viscous ? 5 : 4);
    CFDX_TRY(cfdx::check_launch("block stage"));
    if (b->timing) {
        float t[5] = {0, 0, 0, 0, 0};
        CFDX_HIP(hipEventSynchronize(b->ev[5]), "hipEventSynchronize");
        for (int q = 0; q < 5; ++q) (void)hipEventElapsedTime(&t[q], b->ev[q], b->ev[q + 1]);
        b->t_visc += t[0];
        for (int d = 0; d < 3; ++d) b->t_face[d] += t[1 + d];
        b->t_update += t[4];
    }
    return CFDX_OK;
}

// The captured step is destroyed only after every replay of it has finished on the batch's stream
// (its kernel nodes hold copies of the launch arguments, which the runtime releases with it).
void drop_graph(cfdx_batch *q)
{
    if (q->exec) (void)hipStreamSynchronize(q->ctx->stream);
    if (q->exec) (void)hipGraphExecDestroy(q->exec);
    if (q->graph) (void)hipGraphDestroy(q->graph);
    q->exec = nullptr;
    q->graph = nullptr;
}

int upload_batch(cfdx_batch *q)
{
    bool changed = q->dirty || q->seen.size() != q->blocks.size();
    for (size_t i = 0; !changed && i < q->blocks.size(); ++i) changed = q->seen[i] != q->blocks[i]->version;
    if (!changed) return CFDX_OK;
    const size_t nb = q->blocks.size();
    if (nb > q->cap) {
        if (q->dB) (void)hipFree(q->dB);
        if (q->dprefix) (void)hipFree(q->dprefix);
        q->dB = nullptr;
        q->dprefix = nullptr;
        const size_t cap = std::max<size_t>(nb, 2 * q->cap);
        hipError_t e = hipMalloc((void **)&q->dB, cap * sizeof(BlockDev));
        if (e == hipSuccess) e = hipMalloc((void **)&q->dprefix, 4 * (cap + 1) * sizeof(uint32_t));
        if (e != hipSuccess) { q->cap = 0; return cfdx::fail(CFDX_E_NOMEM, "batch: descriptor allocation failed"); }
        q->cap = cap;
    }
    std::vector<BlockDev> h(nb);
    std::vector<uint32_t> pre(4 * (q->cap + 1), 0);
    uint32_t tot[4] = {0, 0, 0, 0};
    q->any_visc = false;
    for (size_t i = 0; i < nb; ++i) {
        h[i] = q->blocks[i]->B;
        h[i].strip = (uint32_t)q->tiles.strip;
        for (int kind = 0; kind < 4; ++kind) {
            h[i].tile0[kind] = tot[kind];
            pre[kind * (q->cap + 1) + i] = tot[kind];
            tot[kind] += cfdx::blk::tiles_of(q->blocks[i]->B, kind, q->tiles);
        }
        q->any_visc |= h[i].viscous != 0;
    }
    for (int kind = 0; kind < 4; ++kind) pre[kind * (q->cap + 1) + nb] = tot[kind];
    if (nb) CFDX_HIP(hipMemcpyAsync(q->dB, h.data(), nb * sizeof(BlockDev), hipMemcpyHostToDevice, q->ctx->stream), "batch descriptors");
    CFDX_HIP(hipMemcpyAsync(q->dprefix, pre.data(), pre.size() * sizeof(uint32_t), hipMemcpyHostToDevice, q->ctx->stream), "batch prefix");
    CFDX_HIP(hipStreamSynchronize(q->ctx->stream), "batch upload sync");
    q->L.blks = q->dB;
    q->L.nblk = (uint32_t)nb;
    for (int kind = 0; kind < 4; ++kind) {
        q->L.prefix[kind] = q->dprefix + kind * (q->cap + 1);
        q->L.total[kind] = tot[kind];
    }
    q->L.grid_cap = q->grid_cap;
    q->L.tiles = q->tiles;
    q->L.bc_face = 0;
    for (int d = 0; d < 3; ++d) q->L.bc_face_d[d] = 0;
    q->L.any_cut = false;
    q->L.shell_max = 0;
    for (size_t i = 0; i < nb; ++i) bc_extent(h[i], q->L);
    q->seen.resize(nb);
    for (size_t i = 0; i < nb; ++i) q->seen[i] = q->blocks[i]->version;
    q->dirty = false;
    drop_graph(q);
    return CFDX_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// cfdx_block
// ---------------------------------------------------------------------------
extern "C" int64_t cfdx_block_bytes(const cfdx_block_desc *d)
{
    if (!d) return cfdx::fail(CFDX_E_INVALID, "block_bytes: null descriptor");
    const int rc = check_desc(d);
    if (rc != CFDX_OK) return rc;
    return bytes_of(d) + (int64_t)tail_bytes();
}

extern "C" int cfdx_block_padded(const cfdx_block_desc *desc, int32_t dims[3], int32_t off[3])
{
    if (!desc || !dims || !off) return cfdx::fail(CFDX_E_INVALID, "block_padded: null argument");
    int d[3], o[3];
    padded_dims(desc, d, o);
    for (int q = 0; q < 3; ++q) { dims[q] = d[q]; off[q] = o[q]; }
    return CFDX_OK;
}

extern "C" int cfdx_block_create(cfdx_ctx *ctx, const cfdx_block_desc *desc, cfdx_block **out)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (!desc || !out) return cfdx::fail(CFDX_E_INVALID, "block_create: null argument");
    *out = nullptr;
    CFDX_TRY(check_desc(desc));
    const long long n = stored_nodes(desc);
    cfdx_block *b = new cfdx_block;
    b->ctx = ctx;
    cfdx::ctx_retain(ctx);
    b->desc = *desc;
    b->bytes = (size_t)bytes_of(desc);
    b->mem = cfdx::pool_take(ctx->device, b->bytes + tail_bytes(), &b->alloc);
    if (!b->mem) {
        cfdx::ctx_release(ctx);
        delete b;
        return cfdx::fail(CFDX_E_NOMEM, "block_create: no device memory for %zu bytes", b->bytes + tail_bytes());
    }
    b->dB = (BlockDev *)((char *)b->mem + b->bytes);
    b->dprefix = (uint32_t *)((char *)b->dB + ((sizeof(BlockDev) + 255) / 256) * 256);
    hipError_t e = hipMemsetAsync(b->mem, 0, b->bytes, ctx->stream);
    for (auto &ev : b->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        cfdx_block_destroy(b);
        return cfdx::hip_fail(e, "block_create");
    }
    BlockDev &B = b->B;
    B.ni = desc->ni; B.nj = desc->nj; B.nk = desc->nk; B.n = (uint32_t)n;
    {
        int dims[3];
        padded_dims(desc, dims, B.off);
        B.pi = dims[0];
        B.pj = dims[1];
        for (int q = 0; q < 6; ++q) B.bc[q] = desc->bc[q];
        B.cut_i0 = desc->cut_i0;
        B.cut_i1 = desc->cut_i1;
        B.cos_aoa = std::cos(desc->aoa);
        B.sin_aoa = std::sin(desc->aoa);
    }
    char *p = (char *)b->mem;
    auto takeD = [&]() { double *r = (double *)p; p += n * sizeof(double); return r; };
    auto takeS = [&]() { void *r = p; p += n * flux_size(desc->precision); return r; };
    B.det = takeD();
    for (auto &a : B.m) a = takeD();
    for (auto &r : B.kf) for (auto &a : r) a = takeD();
    for (auto &a : B.df) a = takeD();
    for (auto &a : B.U0) a = takeD();
    for (auto &a : B.Us) a = takeD();
    B.mu = takeD();
    for (auto &a : B.P) a = takeS();
    for (auto &r : B.Fv) for (auto &a : r) a = takeS();
    for (auto &r : B.FH) for (auto &a : r) a = takeS();
    B.geo32 = nullptr;
    if (desc->precision == CFDX_MIXED) {
        B.geo32 = (const float *)p;
        p += (size_t)22 * n * sizeof(float);
    }
    B.Lout = nullptr;
    B.strip = 0;
    B.code = nullptr;
    if (desc->periodic != 7) {
        const std::vector<uint8_t> codes = loc_codes(desc);
        B.code = (const uint8_t *)p;
        e = hipMemcpyAsync((void *)B.code, codes.data(), codes.size(), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            cfdx_block_destroy(b);
            return cfdx::hip_fail(e, "block_create codes");
        }
    }
    cfdx_params dp;
    cfdx_params_default(&dp);
    configure(b, 0.5, 0.0, &dp);
    b->configured = false;
    *out = b;
    return CFDX_OK;
}

extern "C" int cfdx_block_destroy(cfdx_block *b)
{
    if (!b) return CFDX_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    // the launches of a batch the block has left, then those of the batches still holding it, end
    // before its memory goes back to the pool
    if (b->left_pending) (void)hipEventSynchronize(b->left);
    for (cfdx_batch *q : b->batches) {
        (void)hipStreamSynchronize(q->ctx->stream);
        auto it = std::find(q->blocks.begin(), q->blocks.end(), b);
        if (it != q->blocks.end()) q->blocks.erase(it);
        q->dirty = true;
    }
    b->batches.clear();
    for (auto &ev : b->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (b->left) (void)hipEventDestroy(b->left);
    if (b->mem) cfdx::pool_give(b->ctx->device, b->mem, b->alloc);   // dB and dprefix live at its end
    if (b->mon) (void)hipFree(b->mon);
    cfdx_ctx *ctx = b->ctx;
    delete b;
    return cfdx::ctx_release(ctx);
}

extern "C" int cfdx_block_set_geometry(cfdx_block *b, const double *det, const cfdx_metric *metric, const cfdx_deter *deter)
{
    if (!b || !det || !metric || !deter) return cfdx::fail(CFDX_E_INVALID, "set_geometry: null argument");
    if (b->desc.periodic != 7)
        return cfdx::fail(CFDX_E_UNSUPPORTED, "set_geometry: a bounded block needs its ghost-node geometry (cfdx_block_set_geometry_soa)");
    CFDX_TRY(enter_block(b));
    const int ni = b->desc.ni, nj = b->desc.nj, nk = b->desc.nk;
    const long long n = (long long)ni * nj * nk;
    // face consistency: the minus face of a node must be the plus face of its -d neighbour
    // (Main.cpp:2368-2435 builds both from the same six node values)
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const long long g = i + (long long)ni * (j + (long long)nj * k);
                const long long nb[3] = {((i + ni - 1) % ni) + (long long)ni * (j + (long long)nj * k),
                                         i + (long long)ni * (((j + nj - 1) % nj) + (long long)nj * k),
                                         i + (long long)ni * (j + (long long)nj * ((k + nk - 1) % nk))};
                const double *dp = &deter[g].ip;
                for (int d = 0; d < 3; ++d) {
                    const double *dm = &deter[nb[d]].ip;
                    const double a = dp[2 * d + 1], c = dm[2 * d];
                    if (std::fabs(a - c) > 1e-13 * std::fabs(c))
                        return cfdx::fail(CFDX_E_INVALID, "set_geometry: half-point det of node %lld dir %d is not face-consistent", g, d);
                    for (int q = 0; q < 3; ++q) {
                        const double x = metric[g].face[2 * d + 1][3 * d + q], y = metric[nb[d]].face[2 * d][3 * d + q];
                        if (std::fabs(x - y) > 1e-13 * std::fabs(y) && std::fabs(x - y) > 1e-15)
                            return cfdx::fail(CFDX_E_INVALID, "set_geometry: face metric of node %lld dir %d is not face-consistent", g, d);
                    }
                }
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
    CFDX_HIP(hipMemcpyAsync((void *)b->B.det, h.data(), h.size() * 8, hipMemcpyHostToDevice, b->ctx->stream), "set_geometry copy");
    cfdx::blk::geometry_finalize(b->B, b->ctx->stream);
    CFDX_TRY(cfdx::check_launch("geometry_finalize"));
    CFDX_HIP(hipStreamSynchronize(b->ctx->stream), "set_geometry sync");
    return CFDX_OK;
}

extern "C" int cfdx_block_set_geometry_soa(cfdx_block *b, const double *geo)
{
    if (!b || !geo) return cfdx::fail(CFDX_E_INVALID, "set_geometry_soa: null argument");
    CFDX_TRY(enter_block(b));
    CFDX_HIP(hipMemcpyAsync((void *)b->B.det, geo, (size_t)22 * b->B.n * 8, hipMemcpyHostToDevice, b->ctx->stream), "set_geometry_soa copy");
    cfdx::blk::geometry_finalize(b->B, b->ctx->stream);
    CFDX_TRY(cfdx::check_launch("geometry_finalize"));
    // finiteness checked on the device; the counter lives in the face-flux scratch
    unsigned *cnt = (unsigned *)b->B.FH[0][0];
    CFDX_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned), b->ctx->stream), "set_geometry_soa clear");
    cfdx::blk::count_nonfinite(b->B.det, (size_t)22 * b->B.n, cnt, b->ctx->stream);
    CFDX_TRY(cfdx::check_launch("k_nonfinite"));
    unsigned bad = 0;
    CFDX_HIP(hipMemcpyAsync(&bad, cnt, sizeof bad, hipMemcpyDeviceToHost, b->ctx->stream), "set_geometry_soa check");
    CFDX_HIP(hipStreamSynchronize(b->ctx->stream), "set_geometry_soa sync");
    if (bad) return cfdx::fail(CFDX_E_NONFINITE, "set_geometry_soa: the geometry holds values that are not finite");
    return CFDX_OK;
}

extern "C" int cfdx_block_set_flow(cfdx_block *b, const cfdx_flow *flow)
{
    if (!b || !flow) return cfdx::fail(CFDX_E_INVALID, "set_flow: null argument");
    CFDX_TRY(enter_block(b));
    const long long n = (long long)b->desc.ni * b->desc.nj * b->desc.nk;
    std::vector<double> h((size_t)5 * n);
    for (long long g = 0; g < n; ++g) {
        h[g] = flow[g].rho; h[n + g] = flow[g].u; h[2 * n + g] = flow[g].v; h[3 * n + g] = flow[g].w; h[4 * n + g] = flow[g].p;
    }
    // stage the double primitives in Us (5 n_owned <= 5 n_stored doubles), then P and U0 = det * Q(P)
    CFDX_HIP(hipMemcpyAsync(b->B.Us[0], h.data(), h.size() * 8, hipMemcpyHostToDevice, b->ctx->stream), "set_flow copy");
    cfdx::blk::base(b->B, b->desc.precision, b->ctx->stream, b->B.Us[0]);
    CFDX_TRY(cfdx::check_launch("k_base"));
    if (b->B.code) {
        // side conditions of the initial state (boundary and ghost nodes, U0 included)
        CFDX_TRY(upload_single(b));
        cfdx::blk::boundary(single_list(b), b->desc.precision, 2, b->ctx->stream);
        CFDX_TRY(cfdx::check_launch("boundary"));
    }
    CFDX_HIP(hipStreamSynchronize(b->ctx->stream), "set_flow sync");
    return CFDX_OK;
}

extern "C" int cfdx_block_get_flow(cfdx_block *b, cfdx_flow *flow)
{
    if (!b || !flow) return cfdx::fail(CFDX_E_INVALID, "get_flow: null argument");
    CFDX_TRY(enter_block(b));
    const long long n = (long long)b->desc.ni * b->desc.nj * b->desc.nk;
    // the viscous node-flux and face-flux arrays (contiguous, 27 flux-type values per stored node
    // >= 72 bytes + 16) are scratch between steps: the records are built there on the device,
    // from a 16-byte aligned start
    cfdx_flow *tmp = (cfdx_flow *)(((uintptr_t)b->B.Fv[0][0] + 15) & ~(uintptr_t)15);
    cfdx::blk::flow_out(b->B, b->desc.precision, b->ctx->stream, tmp, b->last_mach);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(flow, tmp, (size_t)n * sizeof(cfdx_flow), hipMemcpyDeviceToHost, b->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->ctx->stream);
    if (e != hipSuccess) return cfdx::hip_fail(e, "get_flow");
    return CFDX_OK;
}

extern "C" int cfdx_block_init_case(cfdx_block *b, const cfdx_case_desc *c)
{
    if (!b || !c) return cfdx::fail(CFDX_E_INVALID, "init_case: null argument");
    if (!(c->length[0] > 0 && c->length[1] > 0 && c->length[2] > 0) || !(c->mach > 0) || !std::isfinite(c->amp))
        return cfdx::fail(CFDX_E_INVALID, "init_case: lengths and Mach must be positive");
    if (c->init < CFDX_INIT_FREESTREAM || c->init > CFDX_INIT_TGV)
        return cfdx::fail(CFDX_E_INVALID, "init_case: unknown initial state %d", c->init);
    if (b->desc.periodic != 7) return cfdx::fail(CFDX_E_UNSUPPORTED, "init_case: the sinusoidal map is for fully periodic blocks");
    CFDX_TRY(enter_block(b));
    cfdx::blk::init_geometry(b->B, c->length, c->amp, c->phase, b->ctx->stream);
    CFDX_TRY(cfdx::check_launch("init_geometry"));
    cfdx::blk::geometry_finalize(b->B, b->ctx->stream);
    CFDX_TRY(cfdx::check_launch("geometry_finalize"));
    cfdx::blk::init_state(b->B, c->length, c->amp, c->phase, c->init, c->mach, c->beta, c->center[0], c->center[1], b->ctx->stream);
    CFDX_TRY(cfdx::check_launch("init_state"));
    cfdx::blk::base(b->B, b->desc.precision, b->ctx->stream, b->B.Us[0]);
    CFDX_TRY(cfdx::check_launch("k_base"));
    b->last_mach = c->mach;
    return CFDX_OK;
}

extern "C" int cfdx_block_cfl_dt(cfdx_block *b, double cfl, double *dt)
{
    if (!b || !dt || !(cfl > 0)) return cfdx::fail(CFDX_E_INVALID, "cfl_dt: null argument or cfl <= 0");
    CFDX_TRY(enter_block(b));
    // per-workgroup maxima in the face-flux scratch (at most 1025 doubles)
    double *part = (double *)b->B.FH[0][0];
    const int np = cfdx::blk::cfl_parts(b->B);
    cfdx::blk::cfl_radius(b->B, part, b->ctx->stream);
    CFDX_TRY(cfdx::check_launch("cfl_radius"));
    double r = 0.0;
    CFDX_HIP(hipMemcpyAsync(&r, part + np, sizeof(double), hipMemcpyDeviceToHost, b->ctx->stream), "cfl_dt copy");
    CFDX_HIP(hipStreamSynchronize(b->ctx->stream), "cfl_dt sync");
    if (!(r > 0) || !std::isfinite(r)) return cfdx::fail(CFDX_E_NONFINITE, "cfl_dt: spectral radius %g", r);
    *dt = cfl / r;
    return CFDX_OK;
}

extern "C" int cfdx_block_configure(cfdx_block *b, double Mach, double del_t, const cfdx_params *params)
{
    if (!b) return cfdx::fail(CFDX_E_INVALID, "block_configure: null block");
    CFDX_TRY(check_params(params));
    configure(b, Mach, del_t, params);
    return CFDX_OK;
}

extern "C" int cfdx_block_step(cfdx_block *b, double Mach, double del_t, int nsteps, const cfdx_params *params)
{
    if (!b) return cfdx::fail(CFDX_E_INVALID, "block_step: null block");
    CFDX_TRY(check_params(params));
    CFDX_TRY(enter_block(b));
    if (nsteps < 0) return cfdx::fail(CFDX_E_INVALID, "block_step: nsteps < 0");
    configure(b, Mach, del_t, params);
    CFDX_TRY(upload_single(b));
    b->t_visc = b->t_update = 0;
    b->t_face[0] = b->t_face[1] = b->t_face[2] = 0;
    b->launches = 0;
    for (int s = 0; s < nsteps; ++s)
        for (int j = 0; j < 3; ++j) CFDX_TRY(run_stage(b, j));
    b->steps_done += nsteps;
    if (b->check && nsteps > 0) {
        // the RK base holds the state after the step; the counter lives in the face-flux scratch
        unsigned *cnt = (unsigned *)b->B.FH[0][0];
        CFDX_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned), b->ctx->stream), "block_step check clear");
        cfdx::blk::count_nonfinite(b->B.U0[0], (size_t)5 * b->B.n, cnt, b->ctx->stream);
        CFDX_TRY(cfdx::check_launch("k_nonfinite"));
        unsigned bad = 0;
        CFDX_HIP(hipMemcpyAsync(&bad, cnt, sizeof bad, hipMemcpyDeviceToHost, b->ctx->stream), "block_step check");
        CFDX_HIP(hipStreamSynchronize(b->ctx->stream), "block_step check sync");
        if (bad)
            return cfdx::fail(CFDX_E_NONFINITE, "block_step: the state is not finite after step %lld", b->steps_done);
    }
    return CFDX_OK;
}

extern "C" int cfdx_block_check_finite(cfdx_block *b, int32_t *finite)
{
    if (!b || !finite) return cfdx::fail(CFDX_E_INVALID, "check_finite: null argument");
    CFDX_TRY(enter_block(b));
    unsigned *cnt = (unsigned *)b->B.FH[0][0];   // counter in the face-flux scratch
    CFDX_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned), b->ctx->stream), "check_finite clear");
    cfdx::blk::count_nonfinite(b->B.U0[0], (size_t)5 * b->B.n, cnt, b->ctx->stream);
    CFDX_TRY(cfdx::check_launch("k_nonfinite"));
    unsigned bad = 0;
    CFDX_HIP(hipMemcpyAsync(&bad, cnt, sizeof bad, hipMemcpyDeviceToHost, b->ctx->stream), "check_finite copy");
    CFDX_HIP(hipStreamSynchronize(b->ctx->stream), "check_finite sync");
    *finite = bad ? 0 : 1;
    return CFDX_OK;
}

extern "C" int cfdx_block_set_check(cfdx_block *b, int on)
{
    if (!b) return cfdx::fail(CFDX_E_INVALID, "set_check: null block");
    b->check = on != 0;
    return CFDX_OK;
}

extern "C" int cfdx_block_rhs(cfdx_block *b, const cfdx_params *params, double *L_host)
{
    if (!b || !L_host) return cfdx::fail(CFDX_E_INVALID, "block_rhs: null argument");
    CFDX_TRY(check_params(params));
    CFDX_TRY(enter_block(b));
    const long long n = b->B.n;
    double *Ld = nullptr;
    CFDX_HIP(hipMalloc((void **)&Ld, (size_t)5 * n * 8), "block_rhs hipMalloc");
    configure(b, b->last_mach, b->B.del_t, params);
    b->B.Lout = Ld;
    ++b->version;
    int rc = upload_single(b);
    if (rc == CFDX_OK) rc = run_stage(b, 3);
    if (rc == CFDX_OK) {
        std::vector<double> hs((size_t)5 * n);
        hipError_t e = hipMemcpyAsync(hs.data(), Ld, (size_t)5 * n * 8, hipMemcpyDeviceToHost, b->ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->ctx->stream);
        if (e != hipSuccess) rc = cfdx::hip_fail(e, "block_rhs copy");
        const size_t no = (size_t)b->desc.ni * b->desc.nj * b->desc.nk;
        for (int c = 0; c < 5 && rc == CFDX_OK; ++c) gather_owned(b, hs.data() + (size_t)c * n, L_host + c * no);
    }
    b->B.Lout = nullptr;
    ++b->version;
    (void)hipFree(Ld);
    return rc;
}

// ---- right-hand side into device memory, with no allocation ----------------------------------
// The stored-node right-hand side of cfdx_block_rhs_noalloc lives in one fixed array per device
// (64 MiB: 5 values of at most RHS_SCRATCH_NODES stored nodes), held from the library's load to
// the end of the process. Host code never passes the array's name to a kernel: its device
// address is resolved once per device by cfdx::scratch_resolve (from cfdx_ctx_create, with the
// device current) and kept in scratch_dev.
constexpr size_t RHS_SCRATCH_DOUBLES = (size_t)8 << 20;
constexpr long long RHS_SCRATCH_NODES = (long long)(RHS_SCRATCH_DOUBLES / 5);
static __device__ __align__(256) double scratch[RHS_SCRATCH_DOUBLES];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_dev[SCRATCH_MAX_DEVICES] = {};

double *scratch_of(int device)
{
    std::lock_guard<std::mutex> g(scratch_mu);
    return (device >= 0 && device < SCRATCH_MAX_DEVICES) ? scratch_dev[device] : nullptr;
}

// owned-node values (owned order, component-major with stride ni nj nk) from the stored-node
// array Ls (component-major with stride n)
__global__ void k_gather_owned(const double *__restrict__ Ls, long long n, double *__restrict__ Lo, int ni, int nj,
                               int nk, int o0, int o1, int o2, int pi, int pj)
{
    const long long no = (long long)ni * nj * nk;
    const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= no) return;
    const int i = (int)(o % ni);
    const long long r = o / ni;
    const int j = (int)(r % nj), k = (int)(r / nj);
    const size_t s = (size_t)(i + o0) + (size_t)pi * ((size_t)(j + o1) + (size_t)pj * (size_t)(k + o2));
    for (int c = 0; c < 5; ++c) Lo[(size_t)c * no + o] = Ls[(size_t)c * n + s];
}
}  // namespace

int cfdx::scratch_resolve(int device)
{
    if (device < 0 || device >= SCRATCH_MAX_DEVICES)
        return cfdx::fail(CFDX_E_UNSUPPORTED, "device %d: the RHS scratch table holds devices 0..%d", device,
                          SCRATCH_MAX_DEVICES - 1);
    std::lock_guard<std::mutex> g(scratch_mu);
    if (scratch_dev[device]) return CFDX_OK;
    void *p = nullptr;
    CFDX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    scratch_dev[device] = (double *)p;
    return CFDX_OK;
}

extern "C" int cfdx_block_rhs_noalloc(cfdx_block *b, const cfdx_params *params, double *L_dev)
{
    if (!b || !L_dev) return cfdx::fail(CFDX_E_INVALID, "block_rhs_noalloc: null argument");
    CFDX_TRY(check_params(params));
    const long long n = b->B.n;
    if (n > RHS_SCRATCH_NODES)
        return cfdx::fail(CFDX_E_CAPACITY,
                          "block_rhs_noalloc: %lld stored nodes exceed the %lld of the fixed 64 MiB scratch array "
                          "(cfdx_block_rhs has no limit)", n, RHS_SCRATCH_NODES);
    CFDX_TRY(enter_block(b));
    double *Ls = scratch_of(b->ctx->device);
    if (!Ls) {   // not resolved when the context was made: resolve now (the device is current)
        CFDX_TRY(cfdx::scratch_resolve(b->ctx->device));
        Ls = scratch_of(b->ctx->device);
    }
    configure(b, b->last_mach, b->B.del_t, params);
    b->B.Lout = Ls;
    ++b->version;
    int rc = upload_single(b);
    if (rc == CFDX_OK) rc = run_stage(b, 3);
    if (rc == CFDX_OK) {
        const long long no = (long long)b->desc.ni * b->desc.nj * b->desc.nk;
        hipLaunchKernelGGL(k_gather_owned, dim3(cfdx::blocks_for(no, 256)), dim3(256), 0, b->ctx->stream, Ls, n, L_dev,
                           b->desc.ni, b->desc.nj, b->desc.nk, b->B.off[0], b->B.off[1], b->B.off[2], b->B.pi,
                           b->B.pj);
        rc = cfdx::check_launch("gather_owned");
    }
    b->B.Lout = nullptr;
    ++b->version;
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
    if (ms_face) *ms_face = b->t_face[0] + b->t_face[1] + b->t_face[2];
    if (ms_update) *ms_update = b->t_update;
    if (launches) *launches = b->launches;
    return CFDX_OK;
}

extern "C" int cfdx_block_set_monitor(cfdx_block *b, int on)
{
    if (!b) return cfdx::fail(CFDX_E_INVALID, "set_monitor: null block");
    CFDX_TRY(enter_block(b));
    if (on && !b->mon) {
        uint32_t cap = 0;
        for (int v = 0; v < cfdx::blk::N_NODE; ++v) {
            cfdx::blk::TileCfg t;
            t.node = v;
            cap = std::max(cap, mon_triples(b->B, t));
        }
        CFDX_HIP(hipMalloc((void **)&b->mon, (size_t)(cap + 1) * 3 * sizeof(double)), "set_monitor hipMalloc");
        CFDX_HIP(hipMemsetAsync(b->mon, 0, (size_t)(cap + 1) * 3 * sizeof(double), b->ctx->stream), "set_monitor clear");
        b->mon_cap = cap;
        b->mon_count = 0;
    }
    double *want = on ? b->mon : nullptr;
    if (b->B.mon != want) { b->B.mon = want; ++b->version; }
    return CFDX_OK;
}

extern "C" int cfdx_block_residual(cfdx_block *b, double out[3])
{
    if (!b || !out) return cfdx::fail(CFDX_E_INVALID, "residual: null argument");
    if (!b->B.mon) return cfdx::fail(CFDX_E_INVALID, "residual: the monitor is off (cfdx_block_set_monitor)");
    if (b->mon_count == 0) return cfdx::fail(CFDX_E_INVALID, "residual: no step with the monitor on yet");
    CFDX_TRY(enter_block(b));
    double *dout = b->mon + 3 * (size_t)b->mon_cap;
    cfdx::blk::monitor_reduce(b->mon, b->mon_count, dout, b->ctx->stream);
    CFDX_TRY(cfdx::check_launch("monitor_reduce"));
    CFDX_HIP(hipMemcpyAsync(out, dout, 3 * sizeof(double), hipMemcpyDeviceToHost, b->ctx->stream), "residual copy");
    CFDX_HIP(hipStreamSynchronize(b->ctx->stream), "residual sync");
    return CFDX_OK;
}

extern "C" int cfdx_block_timing_split(cfdx_block *b, double ms[5])
{
    if (!b || !ms) return cfdx::fail(CFDX_E_INVALID, "timing_split: null argument");
    ms[0] = b->t_visc;
    for (int d = 0; d < 3; ++d) ms[1 + d] = b->t_face[d];
    ms[4] = b->t_update;
    return CFDX_OK;
}

// ---------------------------------------------------------------------------
// tile variants
// ---------------------------------------------------------------------------
namespace {
int to_cfg(const cfdx_tiles *t, cfdx::blk::TileCfg &c)
{
    if (!t) return cfdx::fail(CFDX_E_INVALID, "tiles: null argument");
    if (t->face[0] < 0 || t->face[0] >= cfdx::blk::N_FACE_I) return cfdx::fail(CFDX_E_INVALID, "tiles: face[0] variant %d out of range", t->face[0]);
    for (int d = 1; d < 3; ++d)
        if (t->face[d] < 0 || t->face[d] >= cfdx::blk::N_FACE_JK) return cfdx::fail(CFDX_E_INVALID, "tiles: face[%d] variant %d out of range", d, t->face[d]);
    if (t->node < 0 || t->node >= cfdx::blk::N_NODE) return cfdx::fail(CFDX_E_INVALID, "tiles: node variant %d out of range", t->node);
    if (t->strip < 0 || t->strip > 4096) return cfdx::fail(CFDX_E_INVALID, "tiles: strip %d out of 0 .. 4096", t->strip);
    for (int d = 0; d < 3; ++d) c.face[d] = t->face[d];
    c.node = t->node;
    c.strip = t->strip;
    return CFDX_OK;
}
}  // namespace

extern "C" int cfdx_tile_variants(int kind, int32_t *count, int32_t *shapes)
{
    if (!count) return cfdx::fail(CFDX_E_INVALID, "tile_variants: null count");
    const int(*tab)[2] = kind == 0 ? cfdx::blk::FACE_I_TILES : kind == 1 ? cfdx::blk::FACE_JK_TILES : cfdx::blk::NODE_TILES;
    const int n = kind == 0 ? cfdx::blk::N_FACE_I : kind == 1 ? cfdx::blk::N_FACE_JK : cfdx::blk::N_NODE;
    if (kind < 0 || kind > 2) return cfdx::fail(CFDX_E_INVALID, "tile_variants: kind %d is not 0 (face i), 1 (face j/k) or 2 (node)", kind);
    *count = n;
    if (shapes)
        for (int v = 0; v < n; ++v) { shapes[2 * v] = tab[v][0]; shapes[2 * v + 1] = tab[v][1]; }
    return CFDX_OK;
}

extern "C" int cfdx_block_set_tiles(cfdx_block *b, const cfdx_tiles *t)
{
    if (!b) return cfdx::fail(CFDX_E_INVALID, "block_set_tiles: null block");
    cfdx::blk::TileCfg c;
    CFDX_TRY(to_cfg(t, c));
    if (!(c == b->tiles)) { b->tiles = c; ++b->version; }
    return CFDX_OK;
}

extern "C" int cfdx_block_get_tiles(cfdx_block *b, cfdx_tiles *t)
{
    if (!b || !t) return cfdx::fail(CFDX_E_INVALID, "block_get_tiles: null argument");
    for (int d = 0; d < 3; ++d) t->face[d] = b->tiles.face[d];
    t->node = b->tiles.node;
    t->strip = b->tiles.strip;
    return CFDX_OK;
}

extern "C" int cfdx_batch_set_tiles(cfdx_batch *q, const cfdx_tiles *t)
{
    if (!q) return cfdx::fail(CFDX_E_INVALID, "batch_set_tiles: null batch");
    cfdx::blk::TileCfg c;
    CFDX_TRY(to_cfg(t, c));
    if (!(c == q->tiles)) { q->tiles = c; q->dirty = true; }
    return CFDX_OK;
}

// ---------------------------------------------------------------------------
// cfdx_batch
// ---------------------------------------------------------------------------
extern "C" int cfdx_batch_create(cfdx_ctx *ctx, int precision, cfdx_batch **out)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (!out) return cfdx::fail(CFDX_E_INVALID, "batch_create: null out");
    if (precision != CFDX_FP64 && precision != CFDX_MIXED) return cfdx::fail(CFDX_E_INVALID, "batch_create: bad precision %d", precision);
    cfdx_batch *q = new cfdx_batch;
    q->ctx = ctx;
    cfdx::ctx_retain(ctx);
    q->precision = precision;
    *out = q;
    return CFDX_OK;
}

extern "C" int cfdx_batch_destroy(cfdx_batch *q)
{
    if (!q) return CFDX_OK;
    (void)hipSetDevice(q->ctx->device);
    (void)hipStreamSynchronize(q->ctx->stream);
    for (cfdx_block *b : q->blocks) {
        auto ib = std::find(b->batches.begin(), b->batches.end(), q);
        if (ib != b->batches.end()) b->batches.erase(ib);
    }
    drop_graph(q);
    if (q->dB) (void)hipFree(q->dB);
    if (q->dprefix) (void)hipFree(q->dprefix);
    cfdx_ctx *ctx = q->ctx;
    delete q;
    return cfdx::ctx_release(ctx);
}

extern "C" int cfdx_batch_add(cfdx_batch *q, cfdx_block *b)
{
    if (!q || !b) return cfdx::fail(CFDX_E_INVALID, "batch_add: null argument");
    if (b->desc.precision != q->precision) return cfdx::fail(CFDX_E_INVALID, "batch_add: block precision differs from the batch");
    if (b->ctx->device != q->ctx->device) return cfdx::fail(CFDX_E_INVALID, "batch_add: block on another device");
    if (!b->configured) return cfdx::fail(CFDX_E_INVALID, "batch_add: configure the block (cfdx_block_configure) first");
    if (std::find(q->blocks.begin(), q->blocks.end(), b) != q->blocks.end()) return cfdx::fail(CFDX_E_INVALID, "batch_add: block already in the batch");
    CFDX_TRY(enter_block(b));
    // the batch's launches start after the work queued on the block's own stream (its set-up)
    if (b->ctx->stream != q->ctx->stream) {
        if (!b->left) CFDX_HIP(hipEventCreateWithFlags(&b->left, hipEventDisableTiming), "batch_add event");
        CFDX_HIP(hipEventRecord(b->left, b->ctx->stream), "batch_add record");
        CFDX_HIP(hipStreamWaitEvent(q->ctx->stream, b->left, 0), "batch_add wait");
        b->left_pending = false;
    }
    q->blocks.push_back(b);
    b->batches.push_back(q);
    q->dirty = true;
    return CFDX_OK;
}

extern "C" int cfdx_batch_remove(cfdx_batch *q, cfdx_block *b)
{
    if (!q || !b) return cfdx::fail(CFDX_E_INVALID, "batch_remove: null argument");
    auto it = std::find(q->blocks.begin(), q->blocks.end(), b);
    if (it == q->blocks.end()) return cfdx::fail(CFDX_E_INVALID, "batch_remove: block not in the batch");
    CFDX_TRY(cfdx::enter(q->ctx));
    if (!b->left) CFDX_HIP(hipEventCreateWithFlags(&b->left, hipEventDisableTiming), "batch_remove event");
    // the batch's launches so far may still read and write the block: its later work waits for them
    CFDX_HIP(hipEventRecord(b->left, q->ctx->stream), "batch_remove record");
    b->left_pending = true;
    q->blocks.erase(it);
    auto ib = std::find(b->batches.begin(), b->batches.end(), q);
    if (ib != b->batches.end()) b->batches.erase(ib);
    q->dirty = true;
    return CFDX_OK;
}

extern "C" int cfdx_batch_size(cfdx_batch *q) { return q ? (int)q->blocks.size() : 0; }

extern "C" int cfdx_batch_set_grid(cfdx_batch *q, uint32_t grid_cap)
{
    if (!q) return cfdx::fail(CFDX_E_INVALID, "batch_set_grid: null batch");
    if (q->grid_cap != grid_cap) { q->grid_cap = grid_cap; q->dirty = true; }
    return CFDX_OK;
}

extern "C" int cfdx_batch_step(cfdx_batch *q, int nsteps, int use_graph)
{
    if (!q) return cfdx::fail(CFDX_E_INVALID, "batch_step: null batch");
    CFDX_TRY(cfdx::enter(q->ctx));
    if (nsteps < 0) return cfdx::fail(CFDX_E_INVALID, "batch_step: nsteps < 0");
    if (q->blocks.empty() || nsteps == 0) return CFDX_OK;
    CFDX_TRY(upload_batch(q));
    for (cfdx_block *b : q->blocks) b->mon_count = mon_triples(b->B, q->tiles);
    hipStream_t st = q->ctx->stream;
    const cfdx::blk::Overlap ov = overlap_of(q->ctx);
    if (!use_graph) {
        for (int s = 0; s < nsteps; ++s)
            for (int j = 0; j < 3; ++j)
                cfdx::blk::stage(q->L, q->precision, j, st, q->any_visc, nullptr, overlap_if(q->ctx, q->precision, ov));
        return cfdx::check_launch("batch stage");
    }
    if (q->exec && q->graph_gen != q->ctx->gen.load()) drop_graph(q);   // overlap switch or stream changed
    if (!q->exec) {
        q->graph_gen = q->ctx->gen.load();
        CFDX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        // the second stream joins the capture through the fork event and is joined back before
        // each update, so the graph holds the overlap too
        for (int j = 0; j < 3; ++j)
            cfdx::blk::stage(q->L, q->precision, j, st, q->any_visc, nullptr, overlap_if(q->ctx, q->precision, ov, true));
        hipError_t e = hipStreamEndCapture(st, &q->graph);
        if (e != hipSuccess) { q->graph = nullptr; return cfdx::hip_fail(e, "hipStreamEndCapture"); }
        e = hipGraphInstantiate(&q->exec, q->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) { drop_graph(q); return cfdx::hip_fail(e, "hipGraphInstantiate"); }
    }
    for (int s = 0; s < nsteps; ++s) CFDX_HIP(hipGraphLaunch(q->exec, st), "hipGraphLaunch");
    return CFDX_OK;
}
