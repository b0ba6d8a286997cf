/*
 * lbm_hip.h -- C-ABI of the MI355X (gfx950) D2Q9 BGK collide-stream library.
 *
 * Plain C: opaque context, plain pointers and sizes, int status returns
 * (0 = ok, negative = error; lbm_hip_last_error() describes the last one).
 * No entry point ever calls exit().
 *
 * The entry points replace the reference's lattice-core and boundary API
 * (IUPU-HPC/memory-aware-LBM lb.h, boundaries.h) on one GPU.  The host-side
 * mirror of that API (include/lb.h, include/boundaries.h) is built on them.
 *
 * Host array layout used by upload/download ("node layout"): the reference's
 * Simulation buffers without the dynamics pointer, i.e. (lx+2)*(ly+2) nodes,
 * node (x, y) at index x*(ly+2)+y (y contiguous, lb.c:89-100), each node
 * holding its 9 populations contiguously in the D2Q9 order of lb.c:51-55.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_ctx lbm_ctx;

/* Status codes. */
enum {
    LBM_OK = 0,
    LBM_ERR_ARG = -1,          /* invalid argument or unsupported configuration */
    LBM_ERR_HIP = -2,          /* a HIP runtime call failed */
    LBM_ERR_NOMEM = -3,        /* device memory exhausted */
    LBM_ERR_UNSUPPORTED = -4,  /* dynamics / variant not available on the device */
    LBM_ERR_STATE = -5         /* call not valid in the context's current state */
};

/* Sweep algorithms (the reference's five, lb.h:99-122).  On the device:
 *   ORIGIN      two kernels per step: collide in place, then stream
 *               (collide + propagate, lb.c:172-341)
 *   COMBINE     one fused pull kernel per step (collide_with_stream,
 *               lb.c:353-437)
 *   TIGHT       the fused kernel advanced two steps per call
 *               (collide_tight, lb.c:442-931)
 *   TIGHT_BLOCK two time steps per launch in LDS-resident 2-D tiles
 *               (collide_tight_block, lb.c:936-1289)
 *   PANEL       two time steps per launch, each wave64 marching a band of
 *               rows with the first step held in registers
 *               (collide_tight_panel_iy, lb.c:1402-1779)
 *   PANEL4      four time steps per launch: the panel march with three levels
 *               held in registers (no counterpart in the reference; the same
 *               per-node operations in the same order).  Uses a third
 *               population buffer; lattices too small for it run as PANEL.
 *   PANEL6      six time steps per launch, five levels in registers (no
 *               counterpart in the reference either).  Uses a third and a
 *               fourth population buffer; lattices too small for it run as
 *               PANEL4.  A remainder of 2 or 3 steps after its six-step
 *               launches is taken as one eight-step launch in place of the
 *               last six (seven levels in registers, a fifth buffer), or as
 *               two four-step launches where the interior is too small.
 * All of them give bit-identical FP64 results, as the reference's serial
 * sweeps do. */
typedef enum {
    LBM_ORIGIN = 0,
    LBM_COMBINE = 1,
    LBM_TIGHT = 2,
    LBM_TIGHT_BLOCK = 3,
    LBM_PANEL = 4,
    LBM_PANEL4 = 5,
    LBM_PANEL6 = 6
} lbm_variant;

/* Population storage.  F64 is the reference's arithmetic.  F32/F16 store the
 * deviations f_i - w_i.  F32 collides in FP64 and rounds each stored value
 * stochastically (keyed by node, collision index and population, so runs are
 * deterministic); F16 collides in FP32 and rounds to nearest.  Wall operators
 * and reductions run in FP64 for both.  No reference counterpart. */
typedef enum { LBM_STORE_F64 = 0, LBM_STORE_F32 = 1, LBM_STORE_F16 = 2 } lbm_storage;

/* Straight-wall operator families and kinds (boundaries.h:72-102). */
typedef enum { LBM_BC_REGULARIZED = 0, LBM_BC_ZOUHE = 1 } lbm_bc_family;
typedef enum { LBM_BC_VELOCITY = 0, LBM_BC_PRESSURE = 1 } lbm_bc_kind;
typedef enum { LBM_SIDE_UPPER = 0, LBM_SIDE_LOWER = 1, LBM_SIDE_LEFT = 2,
               LBM_SIDE_RIGHT = 3 } lbm_side;

/* One side of the channel: every node of row 1 (lower), row ly (upper),
 * column 1 rows 2..ly-1 (left) or column lx rows 2..ly-1 (right) carries
 * this operator (unsteady.c:185-198 assign them in that precedence).
 * p0/p1 are per-position parameters -- (ux, uy) for a velocity operator,
 * (rho, uPar) for a pressure operator (VelocityBCData / PressureBCData,
 * boundaries.h:50-67) -- indexed by x (0..lx+1) on upper/lower sides and by
 * y (0..ly+1) on left/right sides.  NULL means "all zero" (p0) / "all zero"
 * (p1); for a pressure side a NULL p0 means rho = 1. */
typedef struct {
    int kind;          /* lbm_bc_kind */
    int family;        /* lbm_bc_family */
    const double *p0;
    const double *p1;
} lbm_side_bc;

/* The cylinder channel of the reference driver (unsteady.c:84-199): bulk
 * BGK with relaxation omega, bounce-back on the integer disc
 * (x-obst_x)^2 + (y-obst_y)^2 <= obst_r^2, and one straight-wall operator per
 * side.  With zgb != 0 the right side's density is re-extrapolated every step
 * from columns lx-1 and lx-2 (updateZeroGradientBoundary, unsteady.c:205-214)
 * on the device; the right side must then be a pressure side. */
typedef struct {
    int lx, ly;
    int obst_x, obst_y, obst_r;   /* obst_r < 0: no disc */
    double omega;
    lbm_side_bc side[4];          /* indexed by lbm_side */
    int zgb;
    int storage;                  /* lbm_storage */
} lbm_problem;

/* Create a device context for a problem on the current HIP device; allocates
 * the two population buffers (2 x 9 x (lx+2) x (ly+2) populations, padded)
 * and the per-side parameter arrays, from the device's pool when
 * lbm_hip_pool_reserve made one.  Replaces constructSim + setDynamics
 * (lb.c:85-121) for the device.  The populations start at zero, as calloc
 * leaves the reference's buffers.  Waits only for its own stream. */
int lbm_hip_create(const lbm_problem *prob, lbm_ctx **out);
int lbm_hip_destroy(lbm_ctx *ctx);                               /* destructSim, lb.c:109-115 */

/* Device-side initial condition of the reference driver (iniGeometry,
 * unsteady.c:162-199 with iniEquilibrium, lb.c:71-78): every interior node of
 * the lattice set to equilibrium at rho = 1, u = (init_ux[y], 0), the tmp
 * buffer zeroed, the step counter reset. init_ux has ly+2 entries. */
int lbm_hip_init_equilibrium(lbm_ctx *ctx, const double *init_ux);

/* Copy a reference-convention state (Simulation.lattice = post-stream
 * populations, Simulation.tmpLattice = what the last sweep left there) in
 * node layout to / from the device.  Only interior nodes are transferred;
 * download leaves the ghost ring of the host arrays untouched.  outlet_rho
 * (ly+2 entries, may be NULL) is pressureBoundary[y].rho. */
int lbm_hip_upload(lbm_ctx *ctx, const double *lattice, const double *tmp,
                   const double *outlet_rho);
int lbm_hip_download(lbm_ctx *ctx, double *lattice, double *tmp, double *outlet_rho);

/* Advance nsteps time steps with a sweep variant (lb.h:99-122, driver loop
 * unsteady.c:425-531 including the zero-gradient outlet update).  Enqueued on
 * the context's stream; returns without waiting. */
int lbm_hip_step(lbm_ctx *ctx, int variant, int nsteps);

/* The origin sweep split in its two halves, for callers that drive them
 * separately as the reference driver does.  lbm_hip_collide performs one
 * collision (collide / collide_openmp, lb.c:172-243) and needs the streamed
 * state; lbm_hip_propagate performs the streaming (propagate /
 * propagate_openmp, lb.c:246-341) and needs the collided state.  A call in the
 * wrong state returns LBM_ERR_STATE. */
int lbm_hip_collide(lbm_ctx *ctx);
int lbm_hip_propagate(lbm_ctx *ctx);

/* Periodic ghost copy, makePeriodic (lb.c:1783-1817), "to be called after
 * the propagation step": every perimeter node receives, in the slots of the
 * populations that arrive from outside the domain, what left the domain
 * through the opposite side (corners from the opposite corner), exactly as
 * the reference copies them out of the ghost ring its propagate pushed them
 * into.  Acts on the post-stream state; a collided context (after
 * lbm_hip_collide, or a fused or blocked step) is first brought to the
 * post-stream state, as lbm_hip_propagate would, so a following
 * lbm_hip_propagate returns LBM_ERR_STATE.  Every storage format.  Enqueued
 * on the context's stream. */
int lbm_hip_make_periodic(lbm_ctx *ctx);

/* Block-size knobs of the blocked variants (the reference's blk_size /
 * thread_block, unsteady.c:93,258).  TIGHT_BLOCK: tile extent in x and y.
 * PANEL: tile_x = columns per wave's march (>= 3); tile_y is not used.
 * 0 = default.  The results do not depend on either knob. */
int lbm_hip_set_tile(lbm_ctx *ctx, int tile_x, int tile_y);

/* Capture graph_steps consecutive steps into a hipGraph and replay it for
 * long runs: fused steps for combine/tight, graph_steps / 2 two-step
 * launches for panel.  0 launches every step (default).  No counterpart in
 * the reference. */
int lbm_hip_set_graph_steps(lbm_ctx *ctx, int graph_steps);

/* 1 if all work enqueued on the context has finished, 0 if some is still
 * running, negative on error.  Never blocks. */
int lbm_hip_query(lbm_ctx *ctx);

/* Free and total device memory of the current device (hipMemGetInfo). */
int lbm_hip_device_memory(size_t *free_bytes, size_t *total_bytes);

/* A stream-ordered memory pool for the contexts of the current device (no
 * counterpart in the reference; serving support).  After
 * lbm_hip_pool_reserve, contexts created on this device allocate every
 * device buffer from the pool on their own stream and return it there when
 * destroyed, so creating and destroying contexts never waits for other
 * contexts' work (hipMalloc / hipFree do); memory a destroyed context
 * returns stays in the pool for the next one.  bytes > 0 grows the pool by
 * that much up front (0: grow on demand).  Calling it again only grows it.
 * lbm_hip_pool_release destroys the pool once no context uses it
 * (LBM_ERR_STATE while some do); later contexts use hipMalloc again. */
int lbm_hip_pool_reserve(size_t bytes);
int lbm_hip_pool_release(void);

/* Replace the per-position parameters of one side (e.g. a new inlet profile
 * or, with zgb off, prescribed outlet densities). */
int lbm_hip_set_side_params(lbm_ctx *ctx, int side, const double *p0, const double *p1);

/* Macroscopic fields of the current post-stream state (computeMacros,
 * lb.c:134-141, on every interior node): arrays of lx*ly doubles, x slowest
 * (the order of lattice[iX][iY]).  Any pointer may be NULL. */
int lbm_hip_macros(lbm_ctx *ctx, double *rho, double *ux, double *uy);

/* Diagnostics of the current post-stream state, reduced on the device in a
 * fixed order: out[0] drag and out[1] lift on the disc by momentum exchange,
 * out[2] kinetic energy sum(rho |u|^2 / 2) and out[3] mass over non-solid
 * nodes.  No counterpart in the reference. */
int lbm_hip_diagnostics(lbm_ctx *ctx, double out[4]);

/* Steps taken since init/upload, and the state bookkeeping. */
int64_t lbm_hip_steps_done(const lbm_ctx *ctx);

/* Wait for all work of the context. */
int lbm_hip_sync(lbm_ctx *ctx);

/* Time nsteps of a variant with HIP events on the context's stream (work is
 * enqueued exactly as lbm_hip_step does, including the panel sweep's
 * perimeter kernel on its second stream, which is joined back before the
 * closing event).  ms_total receives the elapsed milliseconds.  No
 * counterpart in the reference (its driver times with gettimeofday,
 * unsteady.c:387,558). */
int lbm_hip_time_steps(lbm_ctx *ctx, int variant, int nsteps, double *ms_total);

/* Launch timing of the kernels that carry a sweep's lattice updates
 * (measurement support; no counterpart in the reference, whose driver times
 * its whole loop with gettimeofday, unsteady.c:387,558).  After
 * lbm_hip_kernel_timing(ctx, 1) every launch of such a kernel -- the interior
 * march of PANEL / PANEL4 / PANEL6 (its perimeter kernels run beside it on a
 * second stream and are not included), the tile kernel of TIGHT_BLOCK, the
 * one-step collide kernel of every variant (ORIGIN, COMBINE / TIGHT, the
 * first collision after init or upload and the single-step remainders) -- is
 * bracketed by two HIP events on the stream that kernel runs on.  At most
 * 4096 launches are recorded; launches inside graph captures are not.
 * lbm_hip_kernel_timing(ctx, 0 or 1) also clears the record.
 * lbm_hip_kernel_times waits for the context's work and writes one record
 * per (kernel, steps per launch) in order of first launch: up to max into
 * out, and the number of records into *count. */
typedef struct {
    char kernel[64];   /* kernel name as rocprofv3 prints it, e.g. "k_wpanel_in<double, 2, 6>" */
    int depth;         /* time steps one launch advances: 1, 2, 4 or 6 */
    int launches;      /* recorded launches */
    double ms;         /* their summed durations, milliseconds */
    double min_ms;     /* the shortest and the longest of them */
    double max_ms;
    double nodes;      /* lattice nodes one launch of the kernel updates */
} lbm_kernel_record;
int lbm_hip_kernel_timing(lbm_ctx *ctx, int on);
int lbm_hip_kernel_times(lbm_ctx *ctx, lbm_kernel_record *out, int max, int *count);

/* Device bytes one step of a variant moves by design (populations read plus
 * written), per interior node. */
double lbm_hip_bytes_per_node(const lbm_ctx *ctx, int variant);

/* Total device memory the context holds, in bytes. */
size_t lbm_hip_footprint(const lbm_ctx *ctx);
/* Device memory a problem would need, without allocating: the two
 * population buffers every variant uses. */
size_t lbm_hip_footprint_of(const lbm_problem *prob);
/* The same for stepping it with `variant`: LBM_PANEL4 adds the one further
 * population buffer it allocates on first use, LBM_PANEL6 three (an upper
 * bound: a lattice too small for the deep march runs the shallower sweep and
 * allocates none, and the third only comes with an eight-step launch).  No
 * counterpart in the reference (admission support). */
size_t lbm_hip_footprint_for(const lbm_problem *prob, int variant);

/* Known-answer entry: apply one node operator of the device to n population
 * vectors (9 doubles each, in place, host memory).  which follows the order
 * of boundaries.h:74-102 (0 bounceBack, 1..4 Zou/He velocity upper, lower,
 * left, right, 5..8 Zou/He pressure, 9..12 regularized velocity, 13..16
 * regularized pressure), plus 17 = bgk (lb.c:154-166).  p0/p1 per vector. */
int lbm_hip_apply_node_op(int which, int n, double *f, const double *p0,
                          const double *p1, double omega);

/* lbm_hip_apply_node_op without any device allocation: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * graph capture elsewhere in the process, tight serving loops, harnesses
 * that forbid hipMalloc).  The n vectors and their parameters are staged
 * through a fixed device array of LBM_NODE_OP_MAX x 11 doubles (352 KiB),
 * which every GPU the library runs on holds until the process ends; n above
 * LBM_NODE_OP_MAX returns LBM_ERR_ARG.  Same operators, same kernel, same
 * bits as lbm_hip_apply_node_op.  Synchronous (the null stream).  Calls on
 * one GPU share its array: the library serialises calls from this process,
 * and calls that overlap on one GPU from elsewhere would race. */
#define LBM_NODE_OP_MAX 4096
int lbm_hip_apply_node_op_noalloc(int which, int n, double *f, const double *p0,
                                  const double *p1, double omega);

/* Measurement support (no counterpart in the reference): time a plain
 * device-to-device copy kernel of `bytes` (read once, written once) over
 * iters launches; the HBM yardstick the roofline figures are read against. */
int lbm_hip_copy_bandwidth(size_t bytes, int iters, double *ms_per_copy);

/* ---- batches: many independent channels, one launch per step ----------
 * No counterpart in the reference (one lattice per process there).  All
 * channels must use the same storage format.  A step of the batch is one
 * persistent launch over a table of 256-node tiles of every channel; each
 * node runs the single-lattice node step, so every channel evolves bit for
 * bit as it would in its own context (fused pull variant).  Batches
 * advance two steps per launch by default (lbm_hip_batch_set_depth), with
 * the LDS tiles of the tight_block sweep: the same bits. */
typedef struct lbm_batch lbm_batch;

int lbm_hip_batch_create(const lbm_problem *probs, int n, lbm_batch **out);
int lbm_hip_batch_destroy(lbm_batch *b);
/* init_ux[c] points to ly_c + 2 values (the initial / inlet profile) */
int lbm_hip_batch_init_equilibrium(lbm_batch *b, const double *const *init_ux);
/* Launch shape: workgroups per CU of the persistent launch, the cap on the
 * tile height (32/64/128/256; tiles are 256 nodes), and the number of steps
 * captured per hipGraph (0 = launch eagerly).  0 leaves a knob unchanged
 * except graph_steps. */
int lbm_hip_batch_set_launch(lbm_batch *b, int wgs_per_cu, int ty_cap, int graph_steps);
/* Steps per launch: 1 (the persistent one-step launch above) or 2 (one
 * workgroup per tile of tile_cols columns x 62 rows of every channel, the
 * first step in LDS; needs lx >= 4 in every channel, LBM_ERR_UNSUPPORTED
 * otherwise; the default for such batches).  tile_cols in [3, 33] for FP64
 * (default 8), [3, 69] for FP32 / FP16 (default 14 / 10); 0 keeps it.  Odd
 * remainders and the first step after init_equilibrium run one step. */
int lbm_hip_batch_set_depth(lbm_batch *b, int depth, int tile_cols);
/* The sweep of the multi-step launches: LBM_TIGHT_BLOCK (two steps, the LDS
 * tiles above, the default), LBM_PANEL (two steps: every channel's waves
 * march its columns with the first step in wave64 registers, the panel
 * sweep's own per-wave code over each channel's wave grid of
 * chunk_cols-column chunks, from flat work tables: one launch for the
 * perimeter waves of every channel, on a second stream, beside one for the
 * interior waves), or LBM_PANEL4 / LBM_PANEL6 (four / six steps: the deep
 * register marches over every channel's interior, rows and columns H + 2 ..
 * ly - H - 1 / lx - H - 1 with H = 3 / 5, in chunk_cols-column chunks, one
 * launch for all channels; on the second stream the frame around them in
 * two / three two-step passes through one / two further pool-sized buffers
 * (each pass but the last widened inwards by what the next one reads), and
 * two-step passes of LBM_PANEL for channels too small for the deep march;
 * remainders of two steps take LBM_PANEL).  chunk_cols: even, in [4, 64]
 * (default 6 for LBM_PANEL, 24 for the deep marches; 0 keeps it).  Same bits
 * every way.  LBM_ERR_UNSUPPORTED for a channel the panel sweep refuses
 * (lx < 4, or too tall for its 32-bit offsets). */
int lbm_hip_batch_set_sweep(lbm_batch *b, int variant, int chunk_cols);
int lbm_hip_batch_step(lbm_batch *b, int nsteps);
int lbm_hip_batch_sync(lbm_batch *b);
int lbm_hip_batch_time(lbm_batch *b, int nsteps, double *ms);
/* one channel's state in node layout, as lbm_hip_download */
int lbm_hip_batch_download(lbm_batch *b, int c, double *lattice, double *tmp,
                           double *outlet_rho);
int lbm_hip_batch_macros(lbm_batch *b, int c, double *rho, double *ux, double *uy);
int64_t lbm_hip_batch_cells(const lbm_batch *b);
int64_t lbm_hip_batch_steps_done(const lbm_batch *b);
/* device bytes held: the two population pools, the further buffers of the
 * deep marches once set (one for LBM_PANEL4, two once LBM_PANEL6 was), and
 * the work tables */
size_t lbm_hip_batch_footprint(const lbm_batch *b);

/* Description of the last error on this thread. */
const char *lbm_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* LBM_HIP_H */
