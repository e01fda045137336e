// Structured-block data shared by the block kernels and the block / batch host objects.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ctx.hpp"
#include "physics.hpp"

namespace cfdx {
namespace blk {

// Tile shapes compiled into the kernels; a launch list picks one of each at run time (the
// autotuner times them on the actual grids). All tiles have 256 threads.
//   face sweep along i: (TA lines across = j, TB faces along i)
//   face sweeps along j and k: (TA lines across = i, TB faces along j or k)
//   node kernels: (TX along i, TY along j) at one k
constexpr int N_FACE_I = 3, N_FACE_JK = 4, N_NODE = 4;
constexpr int FACE_I_TILES[N_FACE_I][2] = {{4, 64}, {2, 128}, {8, 32}};
constexpr int FACE_JK_TILES[N_FACE_JK][2] = {{16, 16}, {32, 8}, {64, 4}, {8, 32}};
constexpr int NODE_TILES[N_NODE][2] = {{64, 4}, {32, 8}, {128, 2}, {256, 1}};

struct TileCfg {
    int face[3] = {0, 0, 0};   // variant of the i, j, k face sweeps
    int node = 0;              // variant of the node kernels
    // node tile order: strips of this many tile rows (0: whole planes). Strips of 4 rows by
    // default: a thin block's planes, split over the 8 XCDs, would otherwise give each XCD only a
    // few planes, whose k +- 2 neighbours another XCD reads (C-grid 384x128x32: +1-2 %)
    int strip = 4;
    bool operator==(const TileCfg &o) const
    {
        return face[0] == o.face[0] && face[1] == o.face[1] && face[2] == o.face[2] && node == o.node &&
               strip == o.strip;
    }
};

// Device-side view of one block: sizes, array pointers and per-block run parameters.
// P, Fv and FH hold the flux type (double or float); everything else is double.
struct BlockDev {
    int ni, nj, nk;             // owned nodes
    uint32_t n;                 // stored (padded) nodes: the length of every array
    int pi, pj;                 // padded sizes along i and j (index strides)
    int off[3];                 // ghost layers before the first owned node along i, j, k (3, or 0 if periodic)
    const uint8_t *code;        // bounded blocks: per stored node 0 interior, 1 boundary node, 100 ghost
    int bc[6];                  // side conditions (CFDX_BC_*) of i-, i+, j-, j+, k-, k+
    int cut_i0, cut_i1;         // wall range of a CFDX_BC_WALL_CUT side
    double cos_aoa, sin_aoa;    // far-field flow direction
    uint32_t tile0[4];          // first global tile of this block: node tiles, face tiles x, y, z
    uint32_t strip;             // node tiles: rows of tiles per strip walked plane by plane (0: whole planes)
    const double *det;          // det at the node
    const double *m[9];         // node metrics zeta_x .. xi_z
    const double *kf[3][3];     // face metric row of direction d at the +1/2 face (zeta_?ip, eta_?jp, xi_?kp)
    const double *df[3];        // half-point det at ip, jp, kp
    const float *geo32;         // mixed blocks: float copies of the 22 geometry arrays above, in their
                                // order (det, m, kf, df); null in FP64 blocks
    void *P[5];                 // rho, u, v, w, p used by the flux kernels
    double *U0[5];              // RK base det*Q
    double *Us[5];              // stage vector U1 / U2
    void *Fv[3][4];             // viscous fluxes, components 1..4 (the mass component is identically 0)
    double *mu;
    void *FH[3][5];             // face fluxes F-hat at the +1/2 faces
    double *Lout;               // RHS output (mode 3)
    double *mon;                // residual monitor: per-wave max |du|, |dv|, |de| of the last step (null: off)
    double del_t, mach, eps;
    int viscous;
    cfdx::dev::ViscParams vp;
};

// Thickness of the shell of a bounded block: the loc-selected derivative forms differ from the
// central ones only at a side's boundary node (every direction) and at the node next to it (the
// form that reaches the ghost layers); two nodes from the side the stencil's outer points are the
// boundary node and the interior, which select the central form.
constexpr uint32_t SHELL = 2;

// Nodes of the shell of a bounded block (within SHELL nodes of a side along a bounded direction;
// the node kernels leave them to the shell kernels); 0 for a periodic block.
__host__ __device__ inline uint32_t shell_count(const BlockDev &B)
{
    if (!B.code) return 0;
    const uint32_t ni = B.ni, nj = B.nj, nk = B.nk;
    const uint32_t kb = B.off[2] ? SHELL : 0u, jb = B.off[1] ? SHELL : 0u;
    uint32_t n = 0;
    if (B.off[2]) n += 2 * SHELL * ni * nj;
    if (B.off[1]) n += 2 * SHELL * ni * (nk - 2 * kb);
    if (B.off[0]) n += 2 * SHELL * (nj - 2 * jb) * (nk - 2 * kb);
    return n;
}

// One launch over a list of blocks.
struct LaunchList {
    const BlockDev *blks = nullptr;   // device array
    const uint32_t *prefix[4] = {};   // device arrays of nblk+1 tile offsets (node, face x, y, z)
    uint32_t nblk = 0;
    uint32_t total[4] = {};
    uint32_t grid_cap = 0;            // 0: one workgroup per tile; else persistent grid of this size
    TileCfg tiles;
    uint32_t bc_face = 0;             // largest side face (owned nodes) of the bounded blocks; 0 if none
    uint32_t bc_face_d[3] = {};       // the same per direction (0: no block bounded in that direction)
    bool any_cut = false;             // some block has a wake cut
    uint32_t shell_max = 0;           // largest shell of the bounded blocks
};

// kernels (block_kernels.hip)
// ev (optional): 6 events recorded before the viscous kernel, after it, after each face sweep
// and after the update
// ov (optional, ignored when timing with ev): the viscous node kernel runs on ov->aux, forked
// from st before the face sweeps and joined back before the update, with at most ov->visc_grid
// workgroups, so that it shares the CUs with the (VALU-bound) face sweeps
struct Overlap {
    hipStream_t aux;
    hipEvent_t fork, join;
    uint32_t visc_grid;
};
void stage(const LaunchList &L, int precision, int mode, hipStream_t st, bool any_visc, hipEvent_t *ev,
           const Overlap *ov = nullptr);
// the float copies of the geometry of a mixed block (after every write of its double geometry)
// after every geometry write: face arrays to unit normal and det_f |k|, then the float copy
void geometry_finalize(const BlockDev &B, hipStream_t st);
// P and U0 of the owned nodes from double primitives staged component-major in owned order
void base(const BlockDev &B, int precision, hipStream_t st, const double *stage);
void prims(const BlockDev &B, int precision, hipStream_t st, double *out);
// *count += the number of workgroups that found a value that is not finite in p[0..n)
void count_nonfinite(const double *p, size_t n, unsigned *count, hipStream_t st);
// FLOW records (cfdx_flow) of the owned nodes, owned order, into device memory
void flow_out(const BlockDev &B, int precision, hipStream_t st, cfdx_flow *out, double mach);
uint32_t tiles_of(const BlockDev &B, int kind, const TileCfg &cfg);
// side conditions of the bounded blocks of a list (mode 0..2 after that RK stage, mode 2 also
// refreshing U0), side by side in the order i-, i+, j-, j+, k-, k+; stage() runs it after the update
void boundary(const LaunchList &L, int precision, int mode, hipStream_t st);
// out[0..2] (device) = maxima of count per-wave monitor triples
void monitor_reduce(const double *mon, uint32_t count, double *out, hipStream_t st);

// device-side set-up (init_kernels.hip)
void init_geometry(const BlockDev &B, const double L[3], double amp, const double ph[3], hipStream_t st);
void init_state(const BlockDev &B, const double L[3], double amp, const double ph[3], int kind, double mach, double beta,
                double xc, double yc, hipStream_t st);
int cfl_parts(const BlockDev &B);
void cfl_radius(const BlockDev &B, double *part, hipStream_t st);

}  // namespace blk
}  // namespace cfdx
