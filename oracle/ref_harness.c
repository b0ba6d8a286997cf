/*
 * ref_harness.c -- test infrastructure, NOT product code.
 *
 * A caller for the UNMODIFIED reference sources lb.c and boundaries.c
 * (IUPU-HPC/memory-aware-LBM), compiled from where they lie by
 * oracle/Makefile into oracle/_ref/libref_harness.so.  It plays the role of
 * the reference driver unsteady.c: it defines the globals that lb.c declares
 * extern (lb.c:30-38), builds the cylinder channel with an explicit geometry
 * (unsteady.c:84-199 hard-codes obst_x = lx/4, obst_r = ly/10+1) and runs a
 * chosen sweep variant with the driver's step ordering (unsteady.c:424-531).
 *
 * Only tests/golden/make_golden.py and the CPU tests load this library, to
 * write and cross-check golden vectors.  It is never part of the GPU path.
 */
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include "lb.h"
#include "boundaries.h"

/* ---- globals the reference lb.c expects from its driver (lb.c:30-38) ---- */
PressureBCData *pressureBoundary = NULL;
int blk_size = 1;
double *myrho1 = NULL, *myrho2 = NULL;
int iT = 0, count = 0;
int thread_block = 1;
int NUM_THREADS = 1;
long long global_CM[NUM_EVENTS];

/* ---- driver-side state (the role of unsteady.c:41-82) ---- */
static int h_lx, h_ly;
static double h_omega;
static Simulation h_sim;
static VelocityBCData *h_inlet = NULL;
static Dynamics *h_left = NULL, *h_right = NULL;
static Dynamics h_bulk, h_bounce, h_upper, h_lower;
static VelocityBCData h_wall;

typedef struct {
    int lx, ly;
    int obst_x, obst_y, obst_r;
    double u_max, reynolds;
    int bc_family;   /* 0: regularized (unsteady.c:68-71,135,138); 1: Zou/He */
    int variant;     /* 0 origin, 1 combine, 2 tight, 3 tight_block, 4 panel */
    int openmp;      /* 1: the *_openmp flavour of the variant */
    int nthreads;
    int blk;         /* blk_size, argv[3] of the reference driver */
    int tb;          /* thread_block, argv[4] of the reference driver */
    int steps;       /* STOP */
    int zgb;         /* 1: outlet density by zero-gradient extrapolation */
} ref_cfg;

/* ZGB outlet density update: the reference driver's own routine
 * (unsteady.c:205-214), restated for this harness. */
void updateZeroGradientBoundary(void) {
    double r1, r2, u1, v1, u2, v2;
    for (int y = 2; y <= h_ly - 1; ++y) {
        computeMacros(h_sim.lattice[h_lx - 1][y].fPop, &r1, &u1, &v1);
        computeMacros(h_sim.lattice[h_lx - 2][y].fPop, &r2, &u2, &v2);
        pressureBoundary[y].rho = 4. / 3. * r1 - 1. / 3. * r2;
        pressureBoundary[y].uPar = 0.;
    }
}

static double poiseuille(int y_idx, double u_max) {
    /* unsteady.c:155-159 */
    double yy = (double)(y_idx - 1);
    double L = (double)(h_ly - 1);
    return 4. * u_max / (L * L) * (L * yy - yy * yy);
}

static void release(void) {
    if (h_sim.memoryChunk) destructSim(&h_sim);
    memset(&h_sim, 0, sizeof(h_sim));
    free(h_inlet); h_inlet = NULL;
    free(pressureBoundary); pressureBoundary = NULL;
    free(h_left); h_left = NULL;
    free(h_right); h_right = NULL;
    free(myrho1); myrho1 = NULL;
    free(myrho2); myrho2 = NULL;
}

static void setup(const ref_cfg *c) {
    release();
    h_lx = c->lx; h_ly = c->ly;
    /* unsteady.c:100-103 */
    double nu = c->u_max * 2. * c->obst_r / c->reynolds;
    h_omega = 1. / (3 * nu + 1. / 2.);
    blk_size = c->blk > 0 ? c->blk : 1;
    thread_block = c->tb > 0 ? c->tb : 1;
    NUM_THREADS = c->nthreads > 0 ? c->nthreads : 1;

    h_bulk.dynamicsFun = &bgk; h_bulk.selfData = &h_omega;
    h_bounce.dynamicsFun = &bounceBack; h_bounce.selfData = NULL;
    h_wall.bulkDynamics = &h_bulk; h_wall.ux = 0.; h_wall.uy = 0.;
    h_upper.dynamicsFun = c->bc_family ? &upperZouHe : &upperRegularized;
    h_upper.selfData = &h_wall;
    h_lower.dynamicsFun = c->bc_family ? &lowerZouHe : &lowerRegularized;
    h_lower.selfData = &h_wall;

    /* unsteady.c:113-142 */
    h_inlet = calloc(h_ly + 2, sizeof(VelocityBCData));
    pressureBoundary = calloc(h_ly + 2, sizeof(PressureBCData));
    h_left = calloc(h_ly + 2, sizeof(Dynamics));
    h_right = calloc(h_ly + 2, sizeof(Dynamics));
    myrho1 = calloc(h_ly, sizeof(double));
    myrho2 = calloc(h_ly, sizeof(double));
    for (int y = 2; y <= h_ly - 1; ++y) {
        h_inlet[y].bulkDynamics = &h_bulk;
        h_inlet[y].uy = 0.;
        pressureBoundary[y].bulkDynamics = &h_bulk;
        pressureBoundary[y].rho = 1.;
        pressureBoundary[y].uPar = 0.;
        h_left[y].dynamicsFun = c->bc_family ? &leftZouHe : &leftRegularized;
        h_left[y].selfData = &h_inlet[y];
        h_right[y].dynamicsFun =
            c->bc_family ? &rightPressureZouHe : &rightPressureRegularized;
        h_right[y].selfData = &pressureBoundary[y];
    }

    constructSim(&h_sim, h_lx, h_ly);

    /* unsteady.c:162-199 */
    int ox = c->obst_x, oy = c->obst_y, r = c->obst_r;
    for (int x = 1; x <= h_lx; ++x)
        for (int y = 1; y <= h_ly; ++y) {
            iniEquilibrium(&h_sim.lattice[x][y], 1., poiseuille(y, c->u_max), 0.);
            if ((x - ox) * (x - ox) + (y - oy) * (y - oy) <= r * r)
                setDynamics(&h_sim, x, y, &h_bounce);
            else
                setDynamics(&h_sim, x, y, &h_bulk);
        }
    for (int x = 1; x <= h_lx; ++x) {
        setDynamics(&h_sim, x, 1, &h_lower);
        setDynamics(&h_sim, x, h_ly, &h_upper);
    }
    for (int y = 2; y <= h_ly - 1; ++y) {
        h_inlet[y].ux = poiseuille(y, c->u_max);
        setDynamics(&h_sim, 1, y, &h_left[y]);
        setDynamics(&h_sim, h_lx, y, &h_right[y]);
    }
}

static void copy_out(Node **lat, double *out) {
    /* (lx+2) x (ly+2) x 9, x slowest, population fastest; ghost ring included */
    size_t k = 0;
    for (int x = 0; x < h_lx + 2; ++x)
        for (int y = 0; y < h_ly + 2; ++y)
            for (int i = 0; i < 9; ++i) out[k++] = lat[x][y].fPop[i];
}

double ref_omega(void) { return h_omega; }

/* Run cfg->steps steps.  Outputs (each may be NULL):
 *   lat_out: sim.lattice (post-stream state), (lx+2)(ly+2)*9 doubles
 *   tmp_out: sim.tmpLattice, same shape
 *   prho_out: pressureBoundary[y].rho for y = 0..ly+1
 * Returns the number of steps actually run (two-step variants round up). */
static int run_with(const ref_cfg *c, int periodic_after, double *lat_out, double *tmp_out,
                    double *prho_out);

int ref_run(const ref_cfg *c, double *lat_out, double *tmp_out, double *prho_out) {
    return run_with(c, -1, lat_out, tmp_out, prho_out);
}

/* As ref_run, with the reference's makePeriodic (lb.c:1783-1817) applied to
 * sim.lattice once, right after the step that brings the count to
 * periodic_after (one-step variants only: it is meant to follow a
 * propagation). */
int ref_run_periodic(const ref_cfg *c, int periodic_after, double *lat_out, double *tmp_out,
                     double *prho_out) {
    if (c->variant > 1 || periodic_after < 1) return -1;
    return run_with(c, periodic_after, lat_out, tmp_out, prho_out);
}

/* The reference's makePeriodic on one lattice in the copy_out layout. */
void ref_make_periodic(int lx, int ly, double *lat) {
    Simulation s;
    constructSim(&s, lx, ly);
    size_t k = 0;
    for (int x = 0; x < lx + 2; ++x)
        for (int y = 0; y < ly + 2; ++y)
            for (int i = 0; i < 9; ++i) s.lattice[x][y].fPop[i] = lat[k++];
    makePeriodic(&s);
    k = 0;
    for (int x = 0; x < lx + 2; ++x)
        for (int y = 0; y < ly + 2; ++y)
            for (int i = 0; i < 9; ++i) lat[k++] = s.lattice[x][y].fPop[i];
    destructSim(&s);
}

static int run_with(const ref_cfg *c, int periodic_after, double *lat_out, double *tmp_out,
                    double *prho_out) {
    void (*coll)(Simulation *) = NULL;
    void (*strm)(Simulation *) = NULL;
    int twostep = 0;
    switch (c->variant) {
    case 0: coll = c->openmp ? collide_openmp : collide;
            strm = c->openmp ? propagate_openmp : propagate; break;
    case 1: coll = c->openmp ? collide_with_stream_openmp : collide_with_stream;
            strm = finalize_stream; break;
    case 2: coll = c->openmp ? collide_tight_openmp : collide_tight; twostep = 1; break;
    case 3: coll = c->openmp ? collide_tight_block_openmp : collide_tight_block;
            twostep = 1; break;
    case 4: coll = c->openmp ? collide_tight_panel_iy_openmp : collide_tight_panel_iy;
            twostep = 1; break;
    default: return -1;
    }
#ifdef _OPENMP
    omp_set_num_threads(NUM_THREADS > 0 ? NUM_THREADS : 1);
#endif
    setup(c);
    if (c->openmp) {
#ifdef _OPENMP
        omp_set_num_threads(NUM_THREADS);
#endif
    }
    int ran = 0;
    /* unsteady.c:425-531: ZGB, collision, stream (skipped by two-step sweeps) */
    for (iT = 0; iT < c->steps; iT += twostep ? 2 : 1) {
        if (c->zgb) updateZeroGradientBoundary();
        coll(&h_sim);
        if (!twostep) strm(&h_sim);
        ran += twostep ? 2 : 1;
        if (ran == periodic_after) makePeriodic(&h_sim);
    }
    if (lat_out) copy_out(h_sim.lattice, lat_out);
    if (tmp_out) copy_out(h_sim.tmpLattice, tmp_out);
    if (prho_out)
        for (int y = 0; y < h_ly + 2; ++y) prho_out[y] = pressureBoundary[y].rho;
    return ran;
}

/* CPU baseline timing: the reference's sweep `c->variant` (with its OpenMP
 * flavour when c->openmp) on the channel of `c`, `warmup` untimed steps,
 * then c->steps timed steps in the driver's loop order (unsteady.c:424-531:
 * ZGB, collision, stream), timed with gettimeofday as the driver does
 * (unsteady.c:387,558, eval_tools.h).  *seconds receives the timed loop's
 * wall time.  Returns the number of timed steps run (two-step sweeps round
 * up to even), negative on a bad variant. */
int ref_time(const ref_cfg *c, int warmup, double *seconds) {
    void (*coll)(Simulation *) = NULL;
    void (*strm)(Simulation *) = NULL;
    int twostep = 0;
    switch (c->variant) {
    case 0: coll = c->openmp ? collide_openmp : collide;
            strm = c->openmp ? propagate_openmp : propagate; break;
    case 1: coll = c->openmp ? collide_with_stream_openmp : collide_with_stream;
            strm = finalize_stream; break;
    case 3: coll = c->openmp ? collide_tight_block_openmp : collide_tight_block;
            twostep = 1; break;
    default: return -1;
    }
    setup(c);
#ifdef _OPENMP
    omp_set_num_threads(NUM_THREADS);
#endif
    int ran = 0;
    struct timeval t0, t1;
    for (iT = 0; iT < warmup + c->steps; iT += twostep ? 2 : 1) {
        if (iT == warmup || (iT > warmup && ran == 0)) gettimeofday(&t0, NULL);
        if (c->zgb) updateZeroGradientBoundary();
        coll(&h_sim);
        if (!twostep) strm(&h_sim);
        if (iT >= warmup) ran += twostep ? 2 : 1;
    }
    gettimeofday(&t1, NULL);
    *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-6 * (double)(t1.tv_usec - t0.tv_usec);
    release();
    return ran;
}

/* ---- per-function known-answer entry points ---- */
void ref_macros(const double *f, double *out3) {
    double tmp[9];
    memcpy(tmp, f, sizeof tmp);
    computeMacros(tmp, &out3[0], &out3[1], &out3[2]);
}

double ref_equilibrium(int i, double rho, double ux, double uy, double usqr) {
    return computeEquilibrium(i, rho, ux, uy, usqr);
}

void ref_bgk(double *f, double omega) {
    double w = omega;
    bgk(f, &w);
}

/* which: 0 bounceBack; 1..8 Zou/He (upper, lower, left, right velocity;
 * upper, lower, left, right pressure); 9..16 regularized, same order.
 * p0, p1: (ux, uy) for velocity BCs, (rho, uPar) for pressure BCs. */
int ref_apply_bc(int which, double *f, double p0, double p1, double omega) {
    static void (*const fn[17])(double *, void *) = {
        bounceBack,
        upperZouHe, lowerZouHe, leftZouHe, rightZouHe,
        upperPressureZouHe, lowerPressureZouHe, leftPressureZouHe, rightPressureZouHe,
        upperRegularized, lowerRegularized, leftRegularized, rightRegularized,
        upperPressureRegularized, lowerPressureRegularized, leftPressureRegularized,
        rightPressureRegularized};
    if (which < 0 || which > 16) return -1;
    double w = omega;
    Dynamics bulk = {&bgk, &w};
    VelocityBCData v = {&bulk, p0, p1};
    PressureBCData p = {&bulk, p0, p1};
    void *self = NULL;
    if (which >= 1 && which <= 4) self = &v;
    else if (which >= 5 && which <= 8) self = &p;
    else if (which >= 9 && which <= 12) self = &v;
    else if (which >= 13) self = &p;
    fn[which](f, self);
    return 0;
}

void ref_release(void) { release(); }
