/*
 * lb.c -- host side of include/lb.h: the reference's lattice-core API
 * (IUPU-HPC/memory-aware-LBM lb.h) with its sweeps running on the GPU through
 * the C-ABI of include/lbm_hip.h.
 *
 * Allocation, node initialisation and the small moment/equilibrium helpers
 * are host code (they are the caller's setup and post-processing); every
 * sweep is a device call.  A Simulation is attached to a device context by
 * lbm_bind_device(), which translates its Dynamics table by function
 * identity into the device's per-side operator description.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lb.h"
#include "../../include/boundaries.h"
#include "../../include/lbm_hip.h"

/* ------------------------------------------------------------------ */
/* error text and the Simulation -> device context registry            */
/* ------------------------------------------------------------------ */

static char g_err[512];

static int set_err(int code, const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}

const char *lbm_last_error(void) { return g_err; }

#define MAX_BOUND 64
static struct {
    const Simulation *sim;
    lbm_ctx *ctx;
    int mid_origin;   /* collide() done, propagate() pending */
} g_bound[MAX_BOUND];

static int slot_of(const Simulation *sim) {
    for (int k = 0; k < MAX_BOUND; ++k)
        if (g_bound[k].sim == sim && g_bound[k].ctx) return k;
    return -1;
}

lbm_ctx *lbm_device_of(const Simulation *sim) {
    int k = slot_of(sim);
    return k < 0 ? NULL : g_bound[k].ctx;
}

/* A sweep on an unbound Simulation is a programming error: this library has
 * no CPU sweep to fall back to. */
static int need_device(Simulation *sim, const char *fn) {
    int k = slot_of(sim);
    if (k < 0) {
        fprintf(stderr,
                "%s: Simulation %p is not bound to a GPU; call lbm_bind_device() after "
                "setting up the lattice (this library has no CPU sweeps)\n",
                fn, (void *)sim);
        abort();
    }
    return k;
}

static void check(int rc, const char *fn) {
    if (rc != 0) {
        fprintf(stderr, "%s: device error %d: %s\n", fn, rc, lbm_hip_last_error());
        abort();
    }
}

static void swap_buffers(Simulation *sim) {
    Node **s = sim->lattice;
    sim->lattice = sim->tmpLattice;
    sim->tmpLattice = s;
}

/* ------------------------------------------------------------------ */
/* nodes, allocation, helpers                                          */
/* ------------------------------------------------------------------ */

void constructNode(Node *node) {                       /* lb.c:62-68 */
    memset(node->fPop, 0, sizeof node->fPop);
    node->dynamics = NULL;
}

void iniEquilibrium(Node *node, double rho, double ux, double uy) {   /* lb.c:71-78 */
    double uSqr = ux * ux + uy * uy;
    for (int i = 0; i < 9; ++i) node->fPop[i] = computeEquilibrium(i, rho, ux, uy, uSqr);
}

void constructSim(Simulation *sim, int lx, int ly) {   /* lb.c:85-106 */
    sim->lx = lx;
    sim->ly = ly;
    size_t n = (size_t)(lx + 2) * (ly + 2);
    sim->memoryChunk = calloc(n, sizeof(Node));
    sim->tmpMemoryChunk = calloc(n, sizeof(Node));
    sim->lattice = calloc(lx + 2, sizeof(Node *));
    sim->tmpLattice = calloc(lx + 2, sizeof(Node *));
    if (!sim->memoryChunk || !sim->tmpMemoryChunk || !sim->lattice || !sim->tmpLattice) {
        fprintf(stderr, "constructSim: out of host memory for %dx%d\n", lx, ly);
        abort();
    }
    for (int x = 0; x < lx + 2; ++x) {
        sim->lattice[x] = sim->memoryChunk + (size_t)x * (ly + 2);
        sim->tmpLattice[x] = sim->tmpMemoryChunk + (size_t)x * (ly + 2);
    }
}

void destructSim(Simulation *sim) {                     /* lb.c:109-115 */
    lbm_unbind_device(sim);
    free(sim->lattice);
    free(sim->memoryChunk);
    free(sim->tmpLattice);
    free(sim->tmpMemoryChunk);
}

void setDynamics(Simulation *sim, int iX, int iY, Dynamics *dyn) {   /* lb.c:118-121 */
    sim->lattice[iX][iY].dynamics = dyn;
    sim->tmpLattice[iX][iY].dynamics = dyn;
}

static const int CXH[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CYH[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const double WH[9] = {4. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 9.,
                             1. / 36., 1. / 36., 1. / 36., 1. / 36.};

void computeMacros(double *f, double *rho, double *ux, double *uy) {   /* lb.c:134-141 */
    double north = f[2] + f[5] + f[6];
    double mid = f[0] + f[1] + f[3];
    double south = f[4] + f[7] + f[8];
    *rho = north + mid + south;
    *ux = (f[1] + f[5] + f[8] - (f[3] + f[6] + f[7])) / (*rho);
    *uy = (north - south) / (*rho);
}

double computeEquilibrium(int iPop, double rho, double ux, double uy, double uSqr) {
    double cu = CXH[iPop] * ux + CYH[iPop] * uy;       /* lb.c:144-151 */
    return rho * WH[iPop] * (1. + 3. * cu + 4.5 * cu * cu - 1.5 * uSqr);
}

/* The per-node operators run the device operator on the one node. */
void bgk(double *fPop, void *selfData) {               /* lb.c:154-166 */
    double zero = 0.;
    check(lbm_hip_apply_node_op_noalloc(17, 1, fPop, &zero, &zero, *(double *)selfData), "bgk");
}

/* ------------------------------------------------------------------ */
/* translation of the Dynamics table                                   */
/* ------------------------------------------------------------------ */

typedef void (*dyn_fn)(double *, void *);

/* operator identity -> (side, kind, family); boundaries.h:74-102 */
static int classify_fn(dyn_fn fn, int *side, int *kind, int *family) {
    static const struct { dyn_fn fn; int side, kind, family; } table[] = {
        {upperZouHe, LBM_SIDE_UPPER, LBM_BC_VELOCITY, LBM_BC_ZOUHE},
        {lowerZouHe, LBM_SIDE_LOWER, LBM_BC_VELOCITY, LBM_BC_ZOUHE},
        {leftZouHe, LBM_SIDE_LEFT, LBM_BC_VELOCITY, LBM_BC_ZOUHE},
        {rightZouHe, LBM_SIDE_RIGHT, LBM_BC_VELOCITY, LBM_BC_ZOUHE},
        {upperPressureZouHe, LBM_SIDE_UPPER, LBM_BC_PRESSURE, LBM_BC_ZOUHE},
        {lowerPressureZouHe, LBM_SIDE_LOWER, LBM_BC_PRESSURE, LBM_BC_ZOUHE},
        {leftPressureZouHe, LBM_SIDE_LEFT, LBM_BC_PRESSURE, LBM_BC_ZOUHE},
        {rightPressureZouHe, LBM_SIDE_RIGHT, LBM_BC_PRESSURE, LBM_BC_ZOUHE},
        {upperRegularized, LBM_SIDE_UPPER, LBM_BC_VELOCITY, LBM_BC_REGULARIZED},
        {lowerRegularized, LBM_SIDE_LOWER, LBM_BC_VELOCITY, LBM_BC_REGULARIZED},
        {leftRegularized, LBM_SIDE_LEFT, LBM_BC_VELOCITY, LBM_BC_REGULARIZED},
        {rightRegularized, LBM_SIDE_RIGHT, LBM_BC_VELOCITY, LBM_BC_REGULARIZED},
        {upperPressureRegularized, LBM_SIDE_UPPER, LBM_BC_PRESSURE, LBM_BC_REGULARIZED},
        {lowerPressureRegularized, LBM_SIDE_LOWER, LBM_BC_PRESSURE, LBM_BC_REGULARIZED},
        {leftPressureRegularized, LBM_SIDE_LEFT, LBM_BC_PRESSURE, LBM_BC_REGULARIZED},
        {rightPressureRegularized, LBM_SIDE_RIGHT, LBM_BC_PRESSURE, LBM_BC_REGULARIZED},
    };
    for (size_t k = 0; k < sizeof table / sizeof table[0]; ++k)
        if (table[k].fn == fn) {
            *side = table[k].side;
            *kind = table[k].kind;
            *family = table[k].family;
            return 1;
        }
    return 0;
}

/* side of perimeter node (x, y) in the reference's assignment order
 * (unsteady.c:185-198), or -1 */
static int side_at(int lx, int ly, int x, int y) {
    if (y == 1) return LBM_SIDE_LOWER;
    if (y == ly) return LBM_SIDE_UPPER;
    if (x == 1) return LBM_SIDE_LEFT;
    if (x == lx) return LBM_SIDE_RIGHT;
    return -1;
}

static int bulk_omega(const Dynamics *d, double *omega) {
    if (!d || d->dynamicsFun != bgk || !d->selfData) return 0;
    *omega = *(const double *)d->selfData;
    return 1;
}

typedef struct {
    lbm_problem prob;
    double *p0[4], *p1[4];
} translation;

static void free_translation(translation *t) {
    for (int s = 0; s < 4; ++s) { free(t->p0[s]); free(t->p1[s]); }
}

static int translate(const Simulation *sim, translation *t) {
    int lx = sim->lx, ly = sim->ly;
    Node **L = sim->lattice;
    memset(t, 0, sizeof *t);
    if (lx < 3 || ly < 3) return set_err(LBM_ERR_ARG, "lattice smaller than 3x3");
    double omega = 0.;
    int have_omega = 0;
    int sx0 = 1 << 30, sx1 = -1, sy0 = 1 << 30, sy1 = -1, nsolid = 0;
    /* bulk and solid nodes */
    for (int x = 1; x <= lx; ++x)
        for (int y = 1; y <= ly; ++y) {
            if (side_at(lx, ly, x, y) >= 0) continue;
            const Dynamics *d = L[x][y].dynamics;
            if (!d) return set_err(LBM_ERR_UNSUPPORTED, "node without dynamics");
            if (d->dynamicsFun == bounceBack) {
                ++nsolid;
                if (x < sx0) sx0 = x;
                if (x > sx1) sx1 = x;
                if (y < sy0) sy0 = y;
                if (y > sy1) sy1 = y;
                continue;
            }
            double w;
            if (!bulk_omega(d, &w))
                return set_err(LBM_ERR_UNSUPPORTED, "interior node with dynamics other than "
                                                    "bgk or bounceBack");
            if (have_omega && w != omega)
                return set_err(LBM_ERR_UNSUPPORTED, "bgk nodes with different omega");
            omega = w;
            have_omega = 1;
        }
    /* the solid set must be the integer disc of the reference driver */
    int ox = 0, oy = 0, r = -1;
    if (nsolid) {
        if ((sx1 - sx0) != (sy1 - sy0) || ((sx1 - sx0) & 1))
            return set_err(LBM_ERR_UNSUPPORTED, "bounce-back set is not a disc");
        ox = (sx0 + sx1) / 2;
        oy = (sy0 + sy1) / 2;
        r = (sx1 - sx0) / 2;
        for (int x = 1; x <= lx; ++x)
            for (int y = 1; y <= ly; ++y) {
                if (side_at(lx, ly, x, y) >= 0) continue;
                int in = (x - ox) * (x - ox) + (y - oy) * (y - oy) <= r * r;
                int solid = L[x][y].dynamics->dynamicsFun == bounceBack;
                if (in != solid)
                    return set_err(LBM_ERR_UNSUPPORTED, "bounce-back set is not a disc");
            }
    }
    /* the four sides */
    for (int s = 0; s < 4; ++s) {
        int n = (s == LBM_SIDE_UPPER || s == LBM_SIDE_LOWER) ? lx + 2 : ly + 2;
        t->p0[s] = calloc(n, sizeof(double));
        t->p1[s] = calloc(n, sizeof(double));
        if (!t->p0[s] || !t->p1[s]) { free_translation(t); return set_err(LBM_ERR_NOMEM, "host"); }
        t->prob.side[s].kind = -1;
    }
    for (int x = 1; x <= lx; ++x)
        for (int y = 1; y <= ly; ++y) {
            int s = side_at(lx, ly, x, y);
            if (s < 0) continue;
            const Dynamics *d = L[x][y].dynamics;
            int side, kind, fam;
            if (!d || !classify_fn(d->dynamicsFun, &side, &kind, &fam) || side != s) {
                free_translation(t);
                return set_err(LBM_ERR_UNSUPPORTED,
                               "perimeter node without the straight-wall operator of its side");
            }
            lbm_side_bc *b = &t->prob.side[s];
            if (b->kind >= 0 && (b->kind != kind || b->family != fam)) {
                free_translation(t);
                return set_err(LBM_ERR_UNSUPPORTED, "mixed operators on one side");
            }
            b->kind = kind;
            b->family = fam;
            const Dynamics *bulk;
            int pos = (s == LBM_SIDE_UPPER || s == LBM_SIDE_LOWER) ? x : y;
            if (kind == LBM_BC_VELOCITY) {
                const VelocityBCData *v = d->selfData;
                t->p0[s][pos] = v->ux;
                t->p1[s][pos] = v->uy;
                bulk = v->bulkDynamics;
            } else {
                const PressureBCData *p = d->selfData;
                t->p0[s][pos] = p->rho;
                t->p1[s][pos] = p->uPar;
                bulk = p->bulkDynamics;
            }
            double w;
            if (!bulk_omega(bulk, &w) || (have_omega && w != omega)) {
                free_translation(t);
                return set_err(LBM_ERR_UNSUPPORTED, "wall operator over a different bulk dynamics");
            }
            omega = w;
            have_omega = 1;
        }
    if (!have_omega) { free_translation(t); return set_err(LBM_ERR_UNSUPPORTED, "no bgk node"); }
    t->prob.lx = lx;
    t->prob.ly = ly;
    t->prob.obst_x = ox;
    t->prob.obst_y = oy;
    t->prob.obst_r = r;
    t->prob.omega = omega;
    for (int s = 0; s < 4; ++s) {
        t->prob.side[s].p0 = t->p0[s];
        t->prob.side[s].p1 = t->p1[s];
    }
    return 0;
}

static double *pack(const Simulation *sim, Node **buf) {
    int lx = sim->lx, ly = sim->ly;
    double *a = malloc(sizeof(double) * 9 * (size_t)(lx + 2) * (ly + 2));
    if (!a) return NULL;
    for (int x = 0; x < lx + 2; ++x)
        for (int y = 0; y < ly + 2; ++y)
            memcpy(a + ((size_t)x * (ly + 2) + y) * 9, buf[x][y].fPop, 9 * sizeof(double));
    return a;
}

static void unpack(const Simulation *sim, Node **buf, const double *a) {
    int lx = sim->lx, ly = sim->ly;
    for (int x = 1; x <= lx; ++x)
        for (int y = 1; y <= ly; ++y)
            memcpy(buf[x][y].fPop, a + ((size_t)x * (ly + 2) + y) * 9, 9 * sizeof(double));
}

int lbm_bind_device_ex(Simulation *sim, const lbm_bind_opts *opts) {
    if (!sim) return set_err(LBM_ERR_ARG, "null Simulation");
    if (slot_of(sim) >= 0) return set_err(LBM_ERR_STATE, "Simulation already bound");
    int k = 0;
    while (k < MAX_BOUND && g_bound[k].ctx) ++k;
    if (k == MAX_BOUND) return set_err(LBM_ERR_NOMEM, "too many bound Simulations");
    translation t;
    int rc = translate(sim, &t);
    if (rc) return rc;
    int zgb = opts ? opts->zgb : -1;
    if (zgb < 0) zgb = t.prob.side[LBM_SIDE_RIGHT].kind == LBM_BC_PRESSURE;
    t.prob.zgb = zgb;
    t.prob.storage = opts ? opts->storage : LBM_STORE_F64;
    lbm_ctx *ctx = NULL;
    rc = lbm_hip_create(&t.prob, &ctx);
    double *lat = NULL, *tmp = NULL;
    if (!rc) {
        lat = pack(sim, sim->lattice);
        tmp = pack(sim, sim->tmpLattice);
        if (!lat || !tmp) rc = set_err(LBM_ERR_NOMEM, "host staging");
    } else {
        set_err(rc, lbm_hip_last_error());
    }
    if (!rc) {
        rc = lbm_hip_upload(ctx, lat, tmp, t.p0[LBM_SIDE_RIGHT]);
        if (rc) set_err(rc, lbm_hip_last_error());
    }
    if (!rc && opts && (opts->tile_x || opts->tile_y))
        rc = lbm_hip_set_tile(ctx, opts->tile_x, opts->tile_y);
    free(lat);
    free(tmp);
    free_translation(&t);
    if (rc) {
        if (ctx) lbm_hip_destroy(ctx);
        return rc;
    }
    g_bound[k].sim = sim;
    g_bound[k].ctx = ctx;
    g_bound[k].mid_origin = 0;
    return 0;
}

int lbm_bind_device(Simulation *sim) { return lbm_bind_device_ex(sim, NULL); }

int lbm_sync_host(Simulation *sim) {
    int k = slot_of(sim);
    if (k < 0) return set_err(LBM_ERR_STATE, "Simulation not bound");
    if (g_bound[k].mid_origin)
        return set_err(LBM_ERR_STATE, "between collide() and propagate()");
    size_t n = 9 * (size_t)(sim->lx + 2) * (sim->ly + 2);
    double *lat = calloc(n, sizeof(double)), *tmp = calloc(n, sizeof(double));
    if (!lat || !tmp) { free(lat); free(tmp); return set_err(LBM_ERR_NOMEM, "host staging"); }
    int rc = lbm_hip_download(g_bound[k].ctx, lat, tmp, NULL);
    if (!rc) {
        unpack(sim, sim->lattice, lat);
        unpack(sim, sim->tmpLattice, tmp);
    } else {
        set_err(rc, lbm_hip_last_error());
    }
    free(lat);
    free(tmp);
    return rc;
}

int lbm_unbind_device(Simulation *sim) {
    int k = slot_of(sim);
    if (k < 0) return 0;
    lbm_hip_destroy(g_bound[k].ctx);
    g_bound[k].ctx = NULL;
    g_bound[k].sim = NULL;
    return 0;
}

/* ------------------------------------------------------------------ */
/* sweeps: lb.c:172-1779 on the device                                 */
/* ------------------------------------------------------------------ */

void collide(Simulation *sim) {
    int k = need_device(sim, "collide");
    check(lbm_hip_collide(g_bound[k].ctx), "collide");
    g_bound[k].mid_origin = 1;
}
void collide_openmp(Simulation *sim) { collide(sim); }

void propagate(Simulation *sim) {
    int k = need_device(sim, "propagate");
    check(lbm_hip_propagate(g_bound[k].ctx), "propagate");
    g_bound[k].mid_origin = 0;
    swap_buffers(sim);                                  /* lb.c:269-271 */
}
void propagate_openmp(Simulation *sim) { propagate(sim); }

void finalize_stream(Simulation *sim) { swap_buffers(sim); }   /* lb.c:344-348 */

void collide_with_stream(Simulation *sim) {
    int k = need_device(sim, "collide_with_stream");
    check(lbm_hip_step(g_bound[k].ctx, LBM_COMBINE, 1), "collide_with_stream");
}
void collide_with_stream_openmp(Simulation *sim) { collide_with_stream(sim); }

/* two steps per call, no buffer swap (the driver skips stream_func) */
void collide_tight(Simulation *sim) {
    int k = need_device(sim, "collide_tight");
    check(lbm_hip_step(g_bound[k].ctx, LBM_TIGHT, 2), "collide_tight");
}
void collide_tight2(Simulation *sim) { collide_tight(sim); }
void collide_tight_openmp(Simulation *sim) { collide_tight(sim); }

void collide_tight_block(Simulation *sim) {
    int k = need_device(sim, "collide_tight_block");
    check(lbm_hip_step(g_bound[k].ctx, LBM_TIGHT_BLOCK, 2), "collide_tight_block");
}
void collide_tight_block_openmp(Simulation *sim) { collide_tight_block(sim); }

void collide_tight_panel_iy(Simulation *sim) {
    int k = need_device(sim, "collide_tight_panel_iy");
    check(lbm_hip_step(g_bound[k].ctx, LBM_PANEL, 2), "collide_tight_panel_iy");
}
void collide_tight_panel_ix(Simulation *sim) { collide_tight_panel_iy(sim); }
void collide_tight_panel_iy_openmp(Simulation *sim) { collide_tight_panel_iy(sim); }

/* periodic ghost copy (lb.c:1783-1817), after a propagation: on a bound
 * lattice the device applies it to its post-stream state. */
void makePeriodic(Simulation *sim) {
    int k = slot_of(sim);
    if (k >= 0) {
        if (g_bound[k].mid_origin) {
            fprintf(stderr, "makePeriodic: called between collide() and propagate(); "
                            "it follows the propagation step\n");
            abort();
        }
        check(lbm_hip_make_periodic(g_bound[k].ctx), "makePeriodic");
        return;
    }
    int lx = sim->lx, ly = sim->ly;
    Node **lat = sim->lattice;
    for (int x = 1; x <= lx; ++x) {
        lat[x][ly].fPop[4] = lat[x][0].fPop[4];
        lat[x][ly].fPop[7] = lat[x][0].fPop[7];
        lat[x][ly].fPop[8] = lat[x][0].fPop[8];
        lat[x][1].fPop[2] = lat[x][ly + 1].fPop[2];
        lat[x][1].fPop[5] = lat[x][ly + 1].fPop[5];
        lat[x][1].fPop[6] = lat[x][ly + 1].fPop[6];
    }
    for (int y = 1; y <= ly; ++y) {
        lat[1][y].fPop[1] = lat[lx + 1][y].fPop[1];
        lat[1][y].fPop[5] = lat[lx + 1][y].fPop[5];
        lat[1][y].fPop[8] = lat[lx + 1][y].fPop[8];
        lat[lx][y].fPop[3] = lat[0][y].fPop[3];
        lat[lx][y].fPop[6] = lat[0][y].fPop[6];
        lat[lx][y].fPop[7] = lat[0][y].fPop[7];
    }
    lat[1][1].fPop[5] = lat[lx + 1][ly + 1].fPop[5];
    lat[lx][1].fPop[6] = lat[0][ly + 1].fPop[6];
    lat[lx][ly].fPop[7] = lat[0][0].fPop[7];
    lat[1][ly].fPop[8] = lat[lx + 1][0].fPop[8];
}

/* ------------------------------------------------------------------ */
/* output (lb.c:1822-1866): the reference's ASCII formats              */
/* ------------------------------------------------------------------ */

void saveVel(Simulation *sim, char fName[]) {
    int lx = sim->lx, ly = sim->ly;
    FILE *o = fopen(fName, "w");
    if (!o) { fprintf(stderr, "saveVel: cannot open %s\n", fName); return; }
    int k = slot_of(sim);
    if (k >= 0) {
        size_t n = (size_t)lx * ly;
        double *rho = malloc(n * sizeof(double)), *ux = malloc(n * sizeof(double)),
               *uy = malloc(n * sizeof(double));
        if (!rho || !ux || !uy) { fprintf(stderr, "saveVel: out of memory\n"); abort(); }
        check(lbm_hip_macros(g_bound[k].ctx, rho, ux, uy), "saveVel");
        for (int y = 1; y <= ly; ++y)
            for (int x = 1; x <= lx; ++x) {
                size_t i = (size_t)(x - 1) * ly + (y - 1);
                fprintf(o, "%f,%f,%f\n", ux[i], 1.0, uy[i]);
            }
        free(rho); free(ux); free(uy);
    } else {
        for (int y = 1; y <= ly; ++y)
            for (int x = 1; x <= lx; ++x) {
                double r, u, v;
                computeMacros(sim->lattice[x][y].fPop, &r, &u, &v);
                fprintf(o, "%f,%f,%f\n", u, 1.0, v);
            }
    }
    fclose(o);
}

void saveF(Simulation *sim, int iPop, char fName[]) {
    if (slot_of(sim) >= 0) check(lbm_sync_host(sim), "saveF");
    FILE *o = fopen(fName, "w");
    if (!o) { fprintf(stderr, "saveF: cannot open %s\n", fName); return; }
    for (int y = 1; y <= sim->ly; ++y) {
        for (int x = 1; x <= sim->lx; ++x) fprintf(o, "%f ", sim->lattice[x][y].fPop[iPop]);
        fprintf(o, "\n");
    }
    fclose(o);
}
