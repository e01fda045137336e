/*
 * cfdx — MI355X (gfx950) implementation of the per-RK-stage right-hand side and
 * time update of the high-order curvilinear compressible Navier-Stokes / LES
 * solver (reference kernels U_Q.cl, Viscous.cl, Weno.cl driven by Solver.cpp).
 *
 * C ABI only: plain pointers and sizes, no C++ or torch types. Every function
 * returns 0 on success or a negative CFDX_E_* code; cfdx_last_error() gives
 * the message of the last failure on the calling thread. Nothing here exits
 * the process.
 *
 * Three layers:
 *  1. Drop-in kernels (cfdx_uq / cfdx_viscous / cfdx_weno): the reference's
 *     three OpenCL kernels with the same argument order, packed AoS layouts and
 *     per-node semantics, on caller-owned DEVICE buffers. n_launch is the
 *     reference's global size (g_node+1 or sd_node+1; index 0 is the dummy node).
 *  2. Structured blocks (cfdx_block_*): one i,j,k block resident in HBM in
 *     structure-of-arrays form, stepped by directional sweep kernels.
 *  3. Host geometry (cfdx_geom_*): neighbour tables, node metrics, half-point
 *     metrics and initial states that feed either layer (CPU only).
 */
#ifndef CFDX_H
#define CFDX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Error codes                                                              */
/* ------------------------------------------------------------------------ */
#define CFDX_OK             0
#define CFDX_E_INVALID     -1   /* bad argument (null pointer, size, flag) */
#define CFDX_E_HIP         -2   /* HIP runtime call failed */
#define CFDX_E_NOMEM       -3   /* device allocation failed */
#define CFDX_E_UNSUPPORTED -4   /* valid request the library does not implement */
#define CFDX_E_NONFINITE   -5   /* non-finite state detected by a checked step */
#define CFDX_E_CAPACITY    -6   /* request exceeds a fixed capacity */

const char *cfdx_last_error(void);

/* ------------------------------------------------------------------------ */
/* Packed layouts, byte-identical to functions.h:49-112 of the reference     */
/* ------------------------------------------------------------------------ */

/* FLOW_VARIABLES (functions.h:49-51), 72 bytes */
typedef struct __attribute__((packed)) { double u, v, w, p, t, rho, mu, a, e; } cfdx_flow;
/* Solver_Column_matrix (functions.h:53-55), 40 bytes */
typedef struct __attribute__((packed)) { double N0, N1, N2, N3, N4; } cfdx_vec5;
/* GRID (functions.h:57-61), 56 bytes: N1 = +i, N3 = -i, N0 = +j, N2 = -j, N4 = +k, N5 = -k;
 * loc = 0 interior, 1..52 boundary kinds, 100 ghost (Main.cpp:1511, node_locations.cpp). */
typedef struct __attribute__((packed)) { double x, y, z; int32_t N0, N1, N2, N3, N4, N5, ID, loc; } cfdx_grid;
/* ROE_AVERAGING (functions.h:63-70), 288 bytes: faces ip, jp, kp, im, jm, km, each (u, v, w, p, h, rho) */
typedef struct __attribute__((packed)) { double f[6][6]; } cfdx_roe;
/* TRANSFORMED (functions.h:93-98), 504 bytes: node metrics zeta_x..xi_z, then per face
 * ip, im, jp, jm, kp, km the same nine metrics interpolated to the half point */
typedef struct __attribute__((packed)) { double node[9]; double face[6][9]; } cfdx_metric;
/* DETERM (functions.h:110-112), 48 bytes: half-point det at ip, im, jp, jm, kp, km */
typedef struct __attribute__((packed)) { double ip, im, jp, jm, kp, km; } cfdx_deter;

/* ------------------------------------------------------------------------ */
/* Physical / numerical parameters                                           */
/* ------------------------------------------------------------------------ */
#define CFDX_FP64  0   /* all arithmetic in double */
#define CFDX_MIXED 1   /* fluxes, reconstruction, projection, viscous terms in float;
                          state, RK accumulation, metrics and norms in double */

typedef struct {
    double mach;      /* Viscous.cl:30 — Mach number in T and the SGS heat flux (ref 1.5) */
    double reyl;      /* Viscous.cl:31 — Reynolds number; INFINITY = inviscid stresses (ref 500) */
    double pr;        /* Viscous.cl:32 — Prandtl number (ref 0.72) */
    double free_t;    /* Viscous.cl:33 — Sutherland reference temperature (ref 273) */
    double c_mts;     /* Viscous.cl:490 — mixed-time-scale constant; 0 disables SGS (ref 0.03) */
    double c_t;       /* Viscous.cl:491 — MTS time-scale constant (ref 10) */
    double eps;       /* Weno.cl:59 — WCNS epsilon "10e-8" (ref 1e-7) */
    int32_t precision;/* CFDX_FP64 or CFDX_MIXED (structured blocks only) */
    int32_t viscous;  /* 1 = full RHS; 0 = Euler (only when reyl is infinite and the SGS model is off) */
    int32_t sgs;      /* CFDX_SGS_MTS: the reference's mixed-time-scale model (Viscous.cl:490-558);
                         CFDX_SGS_SMAGORINSKY: SGS stress -2 rho nu_t S_ij, nu_t = (c_s Delta)^2 |S|,
                         Delta = det^(1/3), |S| = sqrt(2 S_ij S_ij), the same SGS heat-flux form as
                         the MTS model (no reference counterpart) */
    int32_t pad_;
    double c_s;       /* Smagorinsky constant (0.17); 0 disables that model */
} cfdx_params;
#define CFDX_SGS_MTS         0
#define CFDX_SGS_SMAGORINSKY 1

/* Fills the reference's constants, precision FP64, viscous on, the MTS model (c_s = 0.17 for the
 * Smagorinsky option). */
void cfdx_params_default(cfdx_params *p);

/* ------------------------------------------------------------------------ */
/* Context: one device + one HIP stream                                      */
/* ------------------------------------------------------------------------ */
typedef struct cfdx_ctx cfdx_ctx;

/* device: HIP ordinal. stream: a hipStream_t to launch on, or NULL for a
 * stream owned by the context. */
int cfdx_ctx_create(int device, void *stream, cfdx_ctx **out);
int cfdx_ctx_destroy(cfdx_ctx *ctx);
int cfdx_ctx_set_stream(cfdx_ctx *ctx, void *stream);
void *cfdx_ctx_stream(cfdx_ctx *ctx);
/* Overlap (no reference counterpart): mixed-precision block and batch stages, and the FP64 stages
 * of a captured batch graph, run the viscous node kernel on a second stream of the context, beside
 * the face sweeps of the same stage (forked from and joined back to the context's stream by events;
 * a captured batch graph holds both). On by
 * default (environment CFDX_OVERLAP=0 turns the default off); timed passes (cfdx_block_set_timing)
 * always run in order. Changing it (or the stream, cfdx_ctx_set_stream) makes the batches of the
 * context capture their step again at their next cfdx_batch_step. */
int cfdx_ctx_set_overlap(cfdx_ctx *ctx, int on);
int cfdx_sync(cfdx_ctx *ctx);

/* Number of HIP devices visible to this process. */
int cfdx_device_count(int *count);

/* Free and total device memory of the context's device (hipMemGetInfo; free includes the block
 * memory pooled for reuse); used to admit work. */
int cfdx_mem_info(cfdx_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes);
/* Frees the block memory the library keeps for reuse on ctx's device (cfdx_block_destroy returns
 * a block's allocation to a per-device pool, bounded by a quarter of the device memory, instead
 * of freeing it: freeing waits for the whole device). cfdx_mem_info counts pooled memory as free. */
int cfdx_pool_trim(cfdx_ctx *ctx);

/* Device memory and copies on the context's device (the drop-in kernels take
 * device pointers; callers may equally use their own hipMalloc'd buffers).
 * Copies are ordered on the context stream and complete before returning. */
int cfdx_device_alloc(cfdx_ctx *ctx, size_t bytes, void **dptr);
int cfdx_device_free(cfdx_ctx *ctx, void *dptr);
int cfdx_copy_to_device(cfdx_ctx *ctx, void *dst, const void *src, size_t bytes);
int cfdx_copy_to_host(cfdx_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Device-to-device copy by a plain streaming kernel (16-byte loads and stores);
 * bytes must be a multiple of 16. Also the reference point for HBM counters. */
int cfdx_device_copy(cfdx_ctx *ctx, void *dst, const void *src, size_t bytes);

/* HIP events on the context stream: record into one of 8 slots, then the
 * elapsed device time between two recorded slots (waits for the later one). */
int cfdx_event_record(cfdx_ctx *ctx, int slot);
int cfdx_event_elapsed(cfdx_ctx *ctx, int slot_a, int slot_b, double *ms);

/* ------------------------------------------------------------------------ */
/* 1. Drop-in kernels (device pointers, reference argument order)           */
/* ------------------------------------------------------------------------ */

/* U_Q.cl:36-38 (arguments set at Solver.cpp:277-282). */
int cfdx_uq(cfdx_ctx *ctx, cfdx_vec5 *U_C, cfdx_vec5 *Q_C, const cfdx_grid *Grid_P, const double *det,
            cfdx_flow *FLOW, cfdx_roe *ROE_AVER, int64_t n_launch);

/* Viscous.cl:23-26 (arguments set at Solver.cpp:284-294). The constants the
 * reference hard-codes come from params (NULL = reference values). */
int cfdx_viscous(cfdx_ctx *ctx, const cfdx_vec5 *U_C, cfdx_vec5 *F_C, cfdx_vec5 *E_C, cfdx_vec5 *G_C,
                 cfdx_vec5 *Fv_C, cfdx_vec5 *Ev_C, cfdx_vec5 *Gv_C, const cfdx_grid *Grid_P,
                 const cfdx_metric *metric, const double *det, cfdx_flow *FLOW, int64_t n_launch,
                 const cfdx_params *params);

/* Weno.cl:36-40 (arguments set at Solver.cpp:296-314; note F, E, G come after Fv, Ev, Gv). */
int cfdx_weno(cfdx_ctx *ctx, int j, double Mach, double del_t, const cfdx_deter *deter, const double *det,
              cfdx_vec5 *U0, cfdx_vec5 *U1, cfdx_vec5 *U2, const cfdx_grid *Grid_P, const cfdx_vec5 *Q_C,
              const cfdx_metric *metric, const cfdx_roe *ROE_AVER, const cfdx_vec5 *Fv, const cfdx_vec5 *Ev,
              const cfdx_vec5 *Gv, const cfdx_vec5 *F, const cfdx_vec5 *E, const cfdx_vec5 *G, cfdx_flow *FLOW,
              int64_t n_launch, const cfdx_params *params);

/* ------------------------------------------------------------------------ */
/* 2. Structured blocks                                                      */
/* ------------------------------------------------------------------------ */
typedef struct cfdx_block cfdx_block;

/* Side conditions of the directions that are not periodic (Flow_Initialisation.cpp face rules):
 *   WALL      no-slip: u = v = w = 0, p and T of the first interior node, rho = 1.4 M^2 p / T; ghost
 *             layer m mirrors interior node m with the velocity reversed (:952-988)
 *   FARFIELD  free stream u = cos(aoa), v = sin(aoa), w = 0, p = 1 / (1.4 M^2), T = 1; ghosts copy
 *             the boundary node (:384-418 with the inflow turned by aoa)
 *   OUTFLOW   u, v, w, p, T of the first interior node, rho = 1.4 M^2 p / T; ghosts copy the
 *             boundary node (:91-136)
 *   EXTRAP    the boundary node copies the first interior node; ghost layer m copies interior node
 *             m (:640-700)
 *   WALL_CUT  (j- side only) WALL for cut_i0 <= i <= cut_i1; elsewhere the C-grid wake cut: node
 *             (i, -m) is node (ni-1-i, m), and row j = 0 of the lower half copies the upper half.
 * Bounded directions carry 3 ghost layers on each side; the Mach number is the block's (Weno) Mach. */
enum { CFDX_BC_WALL = 1, CFDX_BC_FARFIELD = 2, CFDX_BC_OUTFLOW = 3, CFDX_BC_EXTRAP = 4, CFDX_BC_WALL_CUT = 5 };

typedef struct {
    int32_t ni, nj, nk;   /* owned nodes along i (zeta), j (eta), k (xi); owned node (i,j,k) is i + ni*(j + nj*k) */
    int32_t periodic;     /* bit 0: i, bit 1: j, bit 2: k periodic; 7 = fully periodic */
    int32_t precision;    /* CFDX_FP64, or CFDX_MIXED (float fluxes, double RK state and metrics) */
    int32_t bc[6];        /* CFDX_BC_* of the sides i-, i+, j-, j+, k-, k+ (periodic directions: ignored) */
    int32_t cut_i0, cut_i1;  /* wall range of a CFDX_BC_WALL_CUT side */
    double aoa;           /* far-field flow angle in the x-y plane, radians */
} cfdx_block_desc;

/* Stored (padded) sizes of a block: owned nodes plus 3 ghost layers on each side of every bounded
 * direction; padded node (i, j, k) is i + dims[0]*(j + dims[1]*k) and owned node (i, j, k) sits at
 * padded (i + off[0], j + off[1], k + off[2]). */
int cfdx_block_padded(const cfdx_block_desc *desc, int32_t dims[3], int32_t off[3]);

/* Bytes of device memory a block of this shape holds; a negative CFDX_E_* code (message in
 * cfdx_last_error) for a descriptor cfdx_block_create would refuse. */
int64_t cfdx_block_bytes(const cfdx_block_desc *desc);
int cfdx_block_create(cfdx_ctx *ctx, const cfdx_block_desc *desc, cfdx_block **out);
int cfdx_block_destroy(cfdx_block *b);

/* Host arrays in node order (n = ni*nj*nk entries, reference AoS layouts; fully periodic blocks). */
int cfdx_block_set_geometry(cfdx_block *b, const double *det, const cfdx_metric *metric, const cfdx_deter *deter);
/* Geometry of any block as 22 arrays of padded length, one after the other: det, the 9 node
 * metrics (zeta_x .. xi_z), the +1/2-face metric rows of i (zeta_?), j (eta_?) and k (xi_?) at each
 * node's +1/2 face, and the three +1/2-face dets. Ghost nodes need their values too. */
int cfdx_block_set_geometry_soa(cfdx_block *b, const double *geo);
/* Flow of the owned nodes (ni*nj*nk entries in owned order). set_flow applies the side conditions
 * of a bounded block (boundary and ghost nodes), as the reference does at start-up, with the Mach
 * number of the last cfdx_block_configure / cfdx_block_step (0.5 before either). */
int cfdx_block_set_flow(cfdx_block *b, const cfdx_flow *flow);
int cfdx_block_get_flow(cfdx_block *b, cfdx_flow *flow);

/* Device-side set-up of a periodic block (no host arrays): geometry of the seeded sinusoidal map
 * x = s1 + amp sin(2pi s2/L2 + ph0) sin(2pi s3/L3 + ph0) (and cyclic; amp = 0 gives the Cartesian
 * box), exact Jacobian -> det and node metrics, half-point metrics, and the initial state.
 * Asynchronous on the context stream. */
enum { CFDX_INIT_FREESTREAM = 0, CFDX_INIT_VORTEX = 1, CFDX_INIT_TGV = 2 };
typedef struct {
    double length[3];   /* period along i, j, k */
    double amp;         /* map amplitude */
    double phase[3];    /* map phases */
    int32_t init;       /* CFDX_INIT_* */
    int32_t pad_;
    double mach;        /* free-stream Mach (p_inf = 1 / (1.4 M^2)) */
    double beta;        /* vortex strength */
    double center[2];   /* vortex centre (x, y) */
} cfdx_case_desc;
int cfdx_block_init_case(cfdx_block *b, const cfdx_case_desc *c);

/* Stable time step of the current state (no counterpart in the reference, which runs a fixed
 * del_t, Main.cpp:47): dt = cfl / max over nodes of (r_c + r_v) with the convective radius
 * r_c = sum_d (|u . grad xi_d| + c |grad xi_d|) and, for a block configured viscous, the viscous
 * radius r_v = 1.372^2 max(4/3, 1.4/Pr) mu / (rho Re) (sum_d |grad xi_d|)^2 (mu by Sutherland's law
 * at the node temperature, with the Mach, Re, Pr and T_inf of the block's last configure).
 * Synchronous. */
int cfdx_block_cfl_dt(cfdx_block *b, double cfl, double *dt);
/* Non-finite check of the block's current state (no counterpart in the reference): *finite = 1 if
 * every value of the RK base U0 (all stored nodes) is finite, else 0. One device pass; synchronous. */
int cfdx_block_check_finite(cfdx_block *b, int32_t *finite);

/* Run parameters of the block (Weno-argument Mach for t = gamma M^2 p/rho, fixed del_t, physical
 * constants, viscous switch), used by cfdx_block_step and by batched steps. */
int cfdx_block_configure(cfdx_block *b, double Mach, double del_t, const cfdx_params *params);

/* nsteps SSP-RK3 steps with fixed del_t (configures the block with these arguments first).
 * Asynchronous on the context stream. */
int cfdx_block_step(cfdx_block *b, double Mach, double del_t, int nsteps, const cfdx_params *params);

/* The stage-0 right-hand side L (Weno.cl:3078-3082) of the current state, n*5 doubles
 * (component-major: L[c*n + node]) copied to the host. Synchronous. */
int cfdx_block_rhs(cfdx_block *b, const cfdx_params *params, double *L_host);

/* cfdx_block_rhs with no allocation: the same L (same kernels, bitwise) written to DEVICE memory
 * L_dev (5 * ni*nj*nk doubles, component-major over the owned nodes), asynchronous on the context
 * stream (L_dev is ready once the stream reaches it; cfdx_sync). No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under stream capture, serving
 * loops, harnesses that forbid hipMalloc). The stored-node L is kept in a fixed 64 MiB array of
 * each device, so blocks of more than 1,677,721 stored nodes are refused (CFDX_E_CAPACITY). Every
 * call on one device shares that array: calls that overlap on one GPU (other streams, other
 * threads) race. Calls no hipMalloc / hipFree. */
int cfdx_block_rhs_noalloc(cfdx_block *b, const cfdx_params *params, double *L_dev);

/* Device time in ms of the kernels of the last cfdx_block_step, per kernel class,
 * measured with HIP events on the context stream when timing is on. */
int cfdx_block_set_timing(cfdx_block *b, int on);
int cfdx_block_timing(cfdx_block *b, double *ms_visc, double *ms_face, double *ms_update, int *launches);

/* Residual monitor (Solver.cpp:546-573): with the monitor on, every step also records the
 * maxima over the block of |u_new - u_old|, |v_new - v_old| and |e_new - e_old| (e = internal
 * energy p / (0.4 rho)) between the start and the end of the step; cfdx_block_residual returns
 * those of the last step (out[0..2], synchronous). set_monitor(on) allocates the partial
 * maxima on first use. */
int cfdx_block_set_monitor(cfdx_block *b, int on);
/* Non-finite detector (no counterpart in the reference, whose runs carry NaN on): with it on,
 * every cfdx_block_step ends by checking the block's state on the device and returns
 * CFDX_E_NONFINITE if any value is NaN or infinite (one device pass over U0 and a sync). */
int cfdx_block_set_check(cfdx_block *b, int on);
int cfdx_block_residual(cfdx_block *b, double out[3]);

/* Per-kernel device time of the timed steps: ms[0] viscous, ms[1..3] face sweeps i, j, k,
 * ms[4] update. */
int cfdx_block_timing_split(cfdx_block *b, double ms[5]);

/* Tile shapes. Every kernel class has a few compiled tile variants; kind 0 = face sweep along i
 * (shape = lines across j, faces along i), kind 1 = face sweeps along j and k (lines across i,
 * faces along j or k), kind 2 = node kernels (nodes along i, along j). cfdx_tile_variants gives
 * the count and, if shapes is not NULL, 2 * count shape numbers. A block or batch launches the
 * variants its cfdx_tiles names (variants 0 and strips of 4 rows by default); results do not
 * depend on them. */
typedef struct {
    int32_t face[3];   /* variant of the i, j, k face sweeps */
    int32_t node;      /* variant of the node kernels */
    int32_t strip;     /* order of the node tiles: strips of this many tile rows (along j), each
                          walked plane by plane, so that a strip's k +- 2 planes are recent in the
                          L2; 0 = whole planes one after another; 0 .. 4096 */
} cfdx_tiles;
int cfdx_tile_variants(int kind, int32_t *count, int32_t *shapes);
int cfdx_block_set_tiles(cfdx_block *b, const cfdx_tiles *t);
int cfdx_block_get_tiles(cfdx_block *b, cfdx_tiles *t);

/* Batches: blocks of different sizes and run parameters stepped together. Each kernel class is
 * one launch over the tiles of every block; a step can be replayed from a captured hipGraph.
 * A batch references its blocks (it does not own them); all share its precision and device. */
/* Static operation count of the scheme's arithmetic (no counterpart in the reference; for
 * measurement): the device physics run once on a counting number type. out[6 g + c], groups g =
 * 0 one face of a face sweep, 1 the conserved state of a node, 2 a node flux along one metric row,
 * 3 the viscous + MTS node fluxes of a node; counts c = add/sub, mul, div, sqrt (incl. 1/sqrt),
 * reciprocal, cube root (a fused multiply-add counts as one add and one mul). The arithmetic
 * written out in the node kernels themselves is not included (see opcount.hip). */
int cfdx_op_count(cfdx_ctx *ctx, const cfdx_params *params, int64_t out[24]);
/* Device arithmetic helpers evaluated elementwise (no counterpart in the reference; for testing
 * the kernels' positive-operand reciprocal, square root, reciprocal square root and inverse cube
 * root): y[t] = f(x[t]) for host arrays of n <= 2^24 values, fn = 0..3 the double forms
 * (1/x, sqrt x, 1/sqrt x, x^(-1/3)), 4..7 the float forms of the mixed precision (x rounded to
 * float, the result widened). Operands must be positive normal numbers (of float for fn 3, whose
 * first estimate is taken in float, and for 4..7). */
int cfdx_math_eval(cfdx_ctx *ctx, int fn, const double *x, double *y, int64_t n);

typedef struct cfdx_batch cfdx_batch;
int cfdx_batch_create(cfdx_ctx *ctx, int precision, cfdx_batch **out);
int cfdx_batch_destroy(cfdx_batch *q);
int cfdx_batch_add(cfdx_batch *q, cfdx_block *b);      /* the block must be configured */
/* A block leaves a batch by cfdx_batch_remove, or by cfdx_block_destroy (which waits for the
 * batch's stream first); cfdx_batch_destroy leaves its blocks alive. */
int cfdx_batch_remove(cfdx_batch *q, cfdx_block *b);
int cfdx_batch_size(cfdx_batch *q);
/* grid_cap = 0: one workgroup per tile; > 0: persistent launches of at most grid_cap workgroups
 * that loop over the tiles. */
int cfdx_batch_set_grid(cfdx_batch *q, uint32_t grid_cap);
/* nsteps RK3 steps of every block; use_graph = 1 captures one step in a hipGraph on first use
 * (re-captured after the batch changes) and replays it. Asynchronous on the context stream. No
 * other thread may call the library during a call that captures: the HIP runtime invalidates the
 * capture when another thread allocates device memory or creates streams meanwhile. */
int cfdx_batch_step(cfdx_batch *q, int nsteps, int use_graph);
/* Tile variants of the batch's launches (the blocks' own settings apply to single-block steps). */
int cfdx_batch_set_tiles(cfdx_batch *q, const cfdx_tiles *t);

/* ------------------------------------------------------------------------ */
/* Bounded-block geometry (host)                                             */
/* ------------------------------------------------------------------------ */
/* Algebraic C-grid around a NACA 4-digit section (closed trailing edge) extruded in span: i runs
 * from the lower wake exit along the wake to the trailing edge, around the section (lower surface,
 * leading edge, upper surface) and out along the upper wake; j runs from the wall / wake cut to a
 * far-field C of the given radius around the trailing edge (sinh stretching, first spacing `wall`);
 * k is periodic over `span`. Sides: i- / i+ outflow, j- wall on the section and wake cut behind it,
 * j+ far field. */
typedef struct {
    double m, p, t;        /* NACA 4-digit: maximum camber, its chordwise position, thickness (chords) */
    double wake;           /* wake length behind the trailing edge (chords) */
    double radius;         /* far-field radius (chords) */
    double wall;           /* first wall-normal spacing (chords) */
    double span;           /* span of the periodic k direction (chords) */
    int32_t n_wake;        /* nodes on each wake line */
    int32_t pad_;
} cfdx_cgrid;
int cfdx_geom_cgrid_desc(int32_t ni, int32_t nj, int32_t nk, int32_t precision, const cfdx_cgrid *g, double aoa,
                         cfdx_block_desc *out);
/* Coordinates of every stored node (3 per node, padded layout of cfdx_block_padded), ghosts included:
 * ghost rows below the wall mirror the grid through it and below the wake cut are the nodes across. */
int cfdx_geom_cgrid_coords(const cfdx_block_desc *d, const cfdx_cgrid *g, double *xyz);
/* Coordinates of a sinusoidally mapped box of periods L over the stored nodes of any descriptor. */
int cfdx_geom_box_coords(const cfdx_block_desc *d, const double L[3], double amp, const double ph[3], double *xyz);
/* The 22 geometry arrays of cfdx_block_set_geometry_soa from stored-node coordinates: compact
 * 6th-order derivatives along grid lines (Grid_Transformation.cpp:45-153; cyclic in periodic
 * directions of period L[d], through the wake cut as one line), node metrics, half-point metrics. */
int cfdx_geom_block_metrics(const cfdx_block_desc *d, const double *xyz, const double L[3], double *geo);

/* ------------------------------------------------------------------------ */
/* On-disk formats (host)                                                    */
/* ------------------------------------------------------------------------ */
/* Restart files <dir>/restart_file_<slot>.neu, slot 1 or 2 written in turn, one line per node
 * "u v w rho p t e mu iter" tab-separated with `digits` significant digits (the reference writes 6,
 * Tecplot_output_files.cpp:115-127; 17 round-trips doubles). */
int cfdx_write_restart(const char *dir, int32_t slot, int64_t n, const cfdx_flow *flow, int64_t iter, int32_t digits);
/* The slot (1 or 2) of the newer complete restart file of n nodes and its iteration; slot 0 and
 * iter -1 when neither is complete (presolver.cpp:86-181). */
int cfdx_restart_latest(const char *dir, int64_t n, int32_t *slot, int64_t *iter);
/* Reads n nodes of a restart file (fails unless complete); flow.a is left 0. */
int cfdx_read_restart(const char *dir, int32_t slot, int64_t n, cfdx_flow *flow, int64_t *iter);
/* Tecplot ASCII point file of a structured block with the reference's variable list
 * (X Y Z U V W RHO P T e mu loc Proc DUCROS) and FEBRICK cells (Tecplot_output_files.cpp:92-113);
 * loc may be NULL (all 0). */
int cfdx_write_tecplot(const char *path, int32_t ni, int32_t nj, int32_t nk, const double *xyz, const cfdx_flow *flow,
                       const int32_t *loc, double solution_time);

/* ------------------------------------------------------------------------ */
/* 3. Host geometry (CPU)                                                    */
/* ------------------------------------------------------------------------ */

/* Neighbour table of a periodic ni x nj x nk block in the reference numbering:
 * node 0 is the dummy node, node (i,j,k) is 1 + i + ni*(j + nj*k), all loc = 0.
 * Grid_P must hold ni*nj*nk + 1 entries; x,y,z are taken from xyz (3 per node, may be NULL). */
int cfdx_geom_periodic_table(int32_t ni, int32_t nj, int32_t nk, const double *xyz, cfdx_grid *Grid_P);

/* Node coordinates, 3 doubles per node, node order i fastest. */
int cfdx_geom_coords_box(int32_t ni, int32_t nj, int32_t nk, const double length[3], double *xyz);
/* x = s1 + a sin(s2+px) sin(s3+px), y = s2 + a sin(s3+py) sin(s1+py), z = s3 + a sin(s1+pz) sin(s2+pz),
 * s = (i, j, k) * length / n, a = amp. */
int cfdx_geom_coords_sinusoid(int32_t ni, int32_t nj, int32_t nk, const double length[3], double amp,
                              const double phase[3], double *xyz);

/* Node Jacobian d(x,y,z)/d(i,j,k) (9 per node: x_zeta, x_eta, x_xi, y_zeta, ..., z_xi) of the two
 * analytic maps above (computational spacing 1, Main.cpp:35-36). */
int cfdx_geom_jacobian_sinusoid(int32_t ni, int32_t nj, int32_t nk, const double length[3], double amp,
                                const double phase[3], double *jac);
/* Node Jacobian from coordinates by the 6th-order compact scheme of Grid_Transformation.cpp:45-53
 * along periodic grid lines (cyclic tridiagonal solve). */
int cfdx_geom_jacobian_compact_periodic(int32_t ni, int32_t nj, int32_t nk, const double length[3],
                                        const double *xyz, double *jac);
/* det and node metrics from the Jacobian (Grid_Transformation.cpp:447, :1029-1037); the face
 * metrics of metric[] are left untouched. n nodes, entries 0..n-1. */
int cfdx_geom_node_metrics(int64_t n, const double *jac, double *det, cfdx_metric *metric);
/* Half-point metrics and dets, Main.cpp:2368-2435, for nodes 1..n_sd of a reference-numbered table. */
int cfdx_geom_half_metrics(const cfdx_grid *Grid_P, int64_t n_sd, cfdx_metric *metric, const double *det,
                           cfdx_deter *deter);

/* Initial states on reference-numbered nodes 1..n (node 0 gets the free stream).
 * Free stream (Flow_Initialisation.cpp:35-48): u=1, v=w=0, p=1/(gamma M^2), T=1. */
int cfdx_geom_init_freestream(int64_t n, double mach, cfdx_flow *flow);
/* Isentropic vortex of strength beta centred at (xc, yc), advected with u = 1 (in free-stream units). */
int cfdx_geom_init_vortex(int64_t n, const double *xyz, double mach, double beta, double xc, double yc,
                          cfdx_flow *flow);
/* Taylor-Green vortex: u = sin x cos y cos z, v = -cos x sin y cos z, w = 0,
 * p = 1/(gamma M^2) + (cos 2x + cos 2y)(cos 2z + 2)/16, T = 1. */
int cfdx_geom_init_tgv(int64_t n, const double *xyz, double mach, cfdx_flow *flow);

#ifdef __cplusplus
}
#endif
#endif /* CFDX_H */
