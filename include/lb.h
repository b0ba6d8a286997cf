/*
 * lb.h -- host C API of the D2Q9 lattice core, drop-in for the reference's
 * lb.h (IUPU-HPC/memory-aware-LBM lb.h:52-143) on an MI355X.
 *
 * The types keep the reference's layouts (Dynamics lb.h:59-62, Node
 * lb.h:72-75, Simulation lb.h:87-93) and every entry point keeps its name and
 * signature, so a driver written against the reference compiles unchanged.
 * Behind them the sweeps run on the GPU through include/lbm_hip.h:
 *
 *   1. build the Simulation exactly as before (constructSim, setDynamics,
 *      iniEquilibrium on sim->lattice nodes);
 *   2. call lbm_bind_device(sim) once: it translates the Dynamics table by
 *      function identity (bgk, bounceBack and the boundaries.h operators),
 *      uploads both buffers and attaches a device context;
 *   3. call the sweeps as the reference driver does (collide + propagate,
 *      collide_with_stream + finalize_stream, collide_tight, ...);
 *   4. lbm_sync_host(sim) copies the device state back into sim->lattice /
 *      sim->tmpLattice (saveVel / saveF do it themselves).
 *
 * A sweep called on a Simulation that is not bound aborts with a message:
 * there is no CPU implementation of the sweeps in this library.
 */
#ifndef LB_H
#define LB_H

#include <stddef.h>

#define ZGB

#define NUM_EVENTS 4

/* per-node dynamics: collision function and its parameters (lb.h:59-62) */
typedef struct {
    void (*dynamicsFun)(double *fPop, void *selfData);
    void *selfData;
} Dynamics;

/* one lattice node: 9 populations and its dynamics (lb.h:72-75) */
typedef struct {
    double fPop[9];
    Dynamics *dynamics;
} Node;

void constructNode(Node *node);                                         /* lb.c:62-68 */
void iniEquilibrium(Node *node, double rho, double ux, double uy);      /* lb.c:71-78 */

/* the lattice: two (lx+2)x(ly+2) node buffers, y contiguous (lb.h:87-93) */
typedef struct {
    int lx, ly;
    Node *memoryChunk;
    Node *tmpMemoryChunk;
    Node **lattice;
    Node **tmpLattice;
} Simulation;

/* defined by the driver in the reference (unsteady.c:205-214) */
void updateZeroGradientBoundary();

/* sweeps (lb.h:99-122); _openmp names are the same device sweeps */
void collide(Simulation *sim);
void collide_openmp(Simulation *sim);
void propagate(Simulation *sim);
void propagate_openmp(Simulation *sim);
void finalize_stream(Simulation *sim);
void collide_with_stream(Simulation *sim);
void collide_with_stream_openmp(Simulation *sim);
void collide_tight(Simulation *sim);
void collide_tight2(Simulation *sim);
void collide_tight_openmp(Simulation *sim);
void collide_tight_block(Simulation *sim);
void collide_tight_block_openmp(Simulation *sim);
void collide_tight_panel_ix(Simulation *sim);
void collide_tight_panel_iy(Simulation *sim);
void collide_tight_panel_iy_openmp(Simulation *sim);

/* allocation and geometry (lb.h:125-127) */
void constructSim(Simulation *sim, int lx, int ly);
void destructSim(Simulation *sim);
void setDynamics(Simulation *sim, int iX, int iY, Dynamics *dyn);

/* lb.h:129-131 */
void makePeriodic(Simulation *sim);
void saveVel(Simulation *sim, char fName[]);
void saveF(Simulation *sim, int iPop, char fName[]);

/* helpers (lb.h:137-143) */
void computeMacros(double *f, double *rho, double *ux, double *uy);
double computeEquilibrium(int iPop, double rho, double ux, double uy, double uSqr);
void bgk(double *fPop, void *selfData);

/* ---- device binding (no counterpart in the reference) ---- */

/* Options of lbm_bind_device_ex.  storage: 0 FP64 (bit-exact), see
 * lbm_hip.h.  zgb: 1 = re-extrapolate the outlet density on the device every
 * step, as the reference driver's updateZeroGradientBoundary does (default
 * when the right side carries a pressure operator). */
typedef struct {
    int storage;
    int zgb;      /* -1 = automatic */
    int tile_x;   /* blocked-variant tile knobs, 0 = the driver's blk_size */
    int tile_y;
} lbm_bind_opts;

/* Attach a device context to sim (translate, upload).  Returns 0 or a
 * negative lbm_hip status; lbm_last_error() describes a failure. */
int lbm_bind_device(Simulation *sim);
int lbm_bind_device_ex(Simulation *sim, const lbm_bind_opts *opts);
/* Copy the device state into sim->lattice / sim->tmpLattice. */
int lbm_sync_host(Simulation *sim);
/* Detach (after an optional lbm_sync_host). */
int lbm_unbind_device(Simulation *sim);
/* The device context of a bound Simulation, or NULL. */
struct lbm_ctx *lbm_device_of(const Simulation *sim);
const char *lbm_last_error(void);

#endif /* LB_H */
