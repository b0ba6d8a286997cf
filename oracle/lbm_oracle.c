/*
 * lbm_oracle.c -- CPU oracle for the D2Q9 BGK collide-stream path.
 *
 * TEST INFRASTRUCTURE ONLY (see lbm_oracle.h for who may use it and how
 * every function is pinned).  Double precision, built with
 * -O3 -ffp-contract=off so that every operation rounds exactly as the
 * reference's gcc -O3 x86-64 build does (SSE2, no fused multiply-add).
 * The one-step sweeps spread their column loops over OpenMP threads, which
 * changes no result (each node's work is independent); the two-step orders
 * stay serial, as their visiting order is what they restate.
 *
 * State convention (the reference's): buffer L holds the post-stream,
 * pre-collision populations of the current step; buffer T holds what the
 * previous step left there.  Both are (lx+2)*(ly+2) nodes with a ghost ring;
 * node (x, y) sits at index x*(ly+2)+y (lb.c:89-100, y contiguous), each node
 * carrying its 9 populations contiguously.
 */
#include <stdlib.h>
#include <string.h>
#include "lbm_oracle.h"

/* D2Q9 velocity set and weights: lb.c:48-55 (velocity order 0 rest, 1 E,
 * 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE). */
static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const double W[9] = {4. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 9.,
                            1. / 36., 1. / 36., 1. / 36., 1. / 36.};
/* opposite directions: boundaries.c:32 */
static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

/* ------------------------------------------------------------------ */
/* moments and equilibrium                                             */
/* ------------------------------------------------------------------ */

/* lb.c:134-141: density summed by lines of constant c_y, velocities by true
 * division. */
static void macros(const double *f, double *rho, double *ux, double *uy) {
    double north = f[2] + f[5] + f[6];
    double mid = f[0] + f[1] + f[3];
    double south = f[4] + f[7] + f[8];
    double r = north + mid + south;
    *rho = r;
    *ux = (f[1] + f[5] + f[8] - (f[3] + f[6] + f[7])) / r;
    *uy = (north - south) / r;
}

/* lb.c:144-151: feq_i = rho w_i (1 + 3 c.u + 9/2 (c.u)^2 - 3/2 u^2), with
 * c.u formed as cx*ux + cy*uy from integer components. */
static double feq(int i, double rho, double ux, double uy, double usqr) {
    double cu = CX[i] * ux + CY[i] * uy;
    return rho * W[i] * (1. + 3. * cu + 4.5 * cu * cu - 1.5 * usqr);
}

/* lb.c:154-166: relax every population, scaling first and adding the
 * weighted equilibrium second (two roundings per population). */
static void bgk_relax(double *f, double omega) {
    double rho, ux, uy;
    macros(f, &rho, &ux, &uy);
    double usqr = ux * ux + uy * uy;
    for (int i = 0; i < 9; ++i) {
        f[i] *= (1 - omega);
        f[i] += omega * feq(i, rho, ux, uy, usqr);
    }
}

/* boundaries.c:48-59: full-way bounce-back, every population swapped with
 * its opposite (a per-call copy instead of the reference's static buffer). */
static void bounce(double *f) {
    double g[9];
    for (int i = 0; i < 9; ++i) g[i] = f[OPP[i]];
    memcpy(f, g, sizeof g);
}

/* ------------------------------------------------------------------ */
/* straight-wall operators: boundaries.c:61-414                        */
/* ------------------------------------------------------------------ */

enum side { UPPER = 0, LOWER = 1, LEFT = 2, RIGHT = 3 };

/* Sum of the populations known at a wall node: the three with zero normal
 * velocity plus twice the three leaving the domain through the wall
 * (boundaries.c:66-126 brackets them exactly this way). */
static double wall_sum(const double *f, int s) {
    switch (s) {
    case UPPER: return f[0] + f[3] + f[1] + 2 * (f[6] + f[2] + f[5]);
    case LOWER: return f[0] + f[3] + f[1] + 2 * (f[8] + f[4] + f[7]);
    case RIGHT: return f[0] + f[2] + f[4] + 2 * (f[1] + f[5] + f[8]);
    default:    return f[0] + f[2] + f[4] + 2 * (f[3] + f[6] + f[7]);
    }
}

/* density implied by a prescribed normal velocity: upperRho/lowerRho/
 * rightRho/leftRho, boundaries.c:66-118 */
static double wall_rho(const double *f, int s, double un) {
    switch (s) {
    case UPPER: return 1. / (1. + un) * wall_sum(f, s);
    case LOWER: return 1. / (1. - un) * wall_sum(f, s);
    case RIGHT: return 1. / (1. + un) * wall_sum(f, s);
    default:    return 1. / (1. - un) * wall_sum(f, s);
    }
}

/* normal velocity implied by a prescribed density: upperU/lowerU/rightU/
 * leftU, boundaries.c:74-126 */
static double wall_un(const double *f, int s, double rho) {
    switch (s) {
    case UPPER: return -1. + 1. / rho * wall_sum(f, s);
    case LOWER: return 1. - 1. / rho * wall_sum(f, s);
    case RIGHT: return -1. + 1. / rho * wall_sum(f, s);
    default:    return 1. - 1. / rho * wall_sum(f, s);
    }
}

/* Zou/He completion of the three unknown populations: completeUpper/Lower/
 * Right/Left, boundaries.c:131-169 */
static void zh_complete(double *f, int s, double ux, double uy, double rho) {
    switch (s) {
    case UPPER:
        f[7] = f[5] + 0.5 * (f[1] - f[3]) - rho * uy / 6 - rho * ux / 2.;
        f[8] = f[6] + 0.5 * (f[3] - f[1]) - rho * uy / 6 + rho * ux / 2.;
        f[4] = f[2] - 2. / 3. * rho * uy;
        break;
    case LOWER:
        f[6] = f[8] + 0.5 * (f[1] - f[3]) + rho * uy / 6 - rho * ux / 2.;
        f[5] = f[7] + 0.5 * (f[3] - f[1]) + rho * uy / 6 + rho * ux / 2.;
        f[2] = f[4] + 2. / 3. * rho * uy;
        break;
    case RIGHT:
        f[6] = f[8] + 0.5 * (f[4] - f[2]) - rho * ux / 6 + rho * uy / 2.;
        f[7] = f[5] + 0.5 * (f[2] - f[4]) - rho * ux / 6 - rho * uy / 2.;
        f[3] = f[1] - 2. / 3. * rho * ux;
        break;
    default:
        f[5] = f[7] + 0.5 * (f[4] - f[2]) + rho * ux / 6. + rho * uy / 2.;
        f[8] = f[6] + 0.5 * (f[2] - f[4]) + rho * ux / 6. - rho * uy / 2.;
        f[1] = f[3] + 2. / 3. * rho * ux;
        break;
    }
}

/* Non-equilibrium momentum flux from the known populations only:
 * upperNeqPi/lowerNeqPi/rightNeqPi/leftNeqPi, boundaries.c:266-302 */
static void neq_pi(const double *n, int s, double *pxx, double *pyy, double *pxy) {
    switch (s) {
    case UPPER:
        *pxx = n[1] + n[3] + 2. * (n[5] + n[6]);
        *pyy = 2. * (n[6] + n[2] + n[5]);
        *pxy = 2. * (n[5] - n[6]);
        break;
    case LOWER:
        *pxx = n[1] + n[3] + 2. * (n[7] + n[8]);
        *pyy = 2. * (n[4] + n[7] + n[8]);
        *pxy = 2. * (n[7] - n[8]);
        break;
    case RIGHT:
        *pxx = 2. * (n[1] + n[5] + n[8]);
        *pyy = n[2] + n[4] + 2. * (n[5] + n[8]);
        *pxy = 2. * (n[5] - n[8]);
        break;
    default:
        *pxx = 2. * (n[3] + n[6] + n[7]);
        *pyy = n[2] + n[4] + 2. * (n[6] + n[7]);
        *pxy = 2. * (n[7] - n[6]);
        break;
    }
}

/* Regularized reconstruction: splitEqNeq + xNeqPi + regularizedF,
 * boundaries.c:239-262: f_i = feq_i + 9/2 w_i Q_i : Pi_neq. */
static void regularize(double *f, int s, double rho, double ux, double uy) {
    double eq[9], neq[9], pxx, pyy, pxy;
    for (int i = 0; i < 9; ++i) {
        eq[i] = feq(i, rho, ux, uy, ux * ux + uy * uy);
        neq[i] = f[i] - eq[i];
    }
    neq_pi(neq, s, &pxx, &pyy, &pxy);
    for (int i = 0; i < 9; ++i)
        f[i] = eq[i] + 9. / 2. * W[i] *
               ((CX[i] * CX[i] - 1. / 3.) * pxx + (CY[i] * CY[i] - 1. / 3.) * pyy +
                2. * CX[i] * CY[i] * pxy);
}

/* Node kinds of the oracle's dynamics table. */
enum kind {
    K_NONE = 0, K_BULK, K_BOUNCE,
    K_ZH_VEL,   /* + side */
    K_ZH_PRES = K_ZH_VEL + 4,
    K_REG_VEL = K_ZH_PRES + 4,
    K_REG_PRES = K_REG_VEL + 4,
    K_END = K_REG_PRES + 4
};

/* One node's collision: the reference's per-node virtual dispatch
 * (lb.c:125-128) resolved by kind.  p0/p1 are (ux, uy) of velocity
 * operators or (rho, uPar) of pressure operators (boundaries.h:50-67). */
static void collide_node(int k, double *f, double p0, double p1, double omega) {
    if (k == K_BULK) { bgk_relax(f, omega); return; }
    if (k == K_BOUNCE) { bounce(f); return; }
    if (k >= K_ZH_VEL && k < K_ZH_PRES) {            /* boundaries.c:171-201 */
        int s = k - K_ZH_VEL;
        double un = (s == UPPER || s == LOWER) ? p1 : p0;
        zh_complete(f, s, p0, p1, wall_rho(f, s, un));
    } else if (k >= K_ZH_PRES && k < K_REG_VEL) {    /* boundaries.c:203-233 */
        int s = k - K_ZH_PRES;
        double un = wall_un(f, s, p0);
        if (s == UPPER || s == LOWER) zh_complete(f, s, p1, un, p0);
        else zh_complete(f, s, un, p1, p0);
    } else if (k >= K_REG_VEL && k < K_REG_PRES) {   /* boundaries.c:304-358 */
        int s = k - K_REG_VEL;
        double un = (s == UPPER || s == LOWER) ? p1 : p0;
        regularize(f, s, wall_rho(f, s, un), p0, p1);
    } else if (k >= K_REG_PRES && k < K_END) {       /* boundaries.c:360-414 */
        int s = k - K_REG_PRES;
        double un = wall_un(f, s, p0);
        if (s == UPPER || s == LOWER) regularize(f, s, p0, p1, un);
        else regularize(f, s, p0, un, p1);
    } else {
        return;
    }
    bgk_relax(f, omega);  /* every wall operator ends with the bulk dynamics */
}

/* ------------------------------------------------------------------ */
/* the cylinder channel (the reference driver, unsteady.c)             */
/* ------------------------------------------------------------------ */

typedef struct {
    int lx, ly, ny;          /* ny = ly + 2 */
    double omega;
    double *L, *T;           /* lattice, tmpLattice */
    unsigned char *kind;     /* per node */
    double *p0, *p1;         /* per-node operator parameters */
    double *prho;            /* outlet density per row, 0..ly+1 */
    double *rho1, *rho2;     /* two-step ZGB captures, 0..ly-1 */
    int zgb;
} chan;

static inline double *node(double *buf, const chan *c, int x, int y) {
    return buf + ((size_t)x * c->ny + y) * 9;
}

double orc_omega_of(double u_max, int obst_r, double reynolds) {
    double nu = u_max * 2. * obst_r / reynolds;      /* unsteady.c:102 */
    return 1. / (3 * nu + 1. / 2.);                 /* unsteady.c:103 */
}

double orc_poiseuille(int y, int ly, double u_max) {
    double yy = (double)(y - 1);                    /* unsteady.c:155-159 */
    double L = (double)(ly - 1);
    return 4. * u_max / (L * L) * (L * yy - yy * yy);
}

static void node_params(chan *c, int x, int y) {
    size_t n = (size_t)x * c->ny + y;
    int k = c->kind[n];
    if (k >= K_ZH_PRES && k < K_REG_VEL && (k - K_ZH_PRES) == RIGHT) c->p0[n] = c->prho[y];
    if (k >= K_REG_PRES && k < K_END && (k - K_REG_PRES) == RIGHT) c->p0[n] = c->prho[y];
}

static void apply(chan *c, double *buf, int x, int y) {
    node_params(c, x, y);
    size_t n = (size_t)x * c->ny + y;
    collide_node(c->kind[n], node(buf, c, x, y), c->p0[n], c->p1[n], c->omega);
}

/* push a node's populations into the other buffer (lb.c:258-267) */
static void push(chan *c, double *src, double *dst, int x, int y) {
    const double *f = node(src, c, x, y);
    for (int i = 0; i < 9; ++i) node(dst, c, x + CX[i], y + CY[i])[i] = f[i];
}

static void chan_free(chan *c) {
    free(c->L); free(c->T); free(c->kind); free(c->p0); free(c->p1);
    free(c->prho); free(c->rho1); free(c->rho2);
    memset(c, 0, sizeof *c);
}

/* unsteady.c:113-142 and 162-199 with an explicit geometry */
static int chan_init(chan *c, const orc_cfg *cfg) {
    memset(c, 0, sizeof *c);
    c->lx = cfg->lx; c->ly = cfg->ly; c->ny = cfg->ly + 2;
    c->zgb = cfg->zgb;
    c->omega = orc_omega_of(cfg->u_max, cfg->obst_r, cfg->reynolds);
    size_t nn = (size_t)(c->lx + 2) * c->ny;
    c->L = calloc(nn * 9, sizeof(double));
    c->T = calloc(nn * 9, sizeof(double));
    c->kind = calloc(nn, 1);
    c->p0 = calloc(nn, sizeof(double));
    c->p1 = calloc(nn, sizeof(double));
    c->prho = calloc(c->ly + 2, sizeof(double));
    c->rho1 = calloc(c->ly, sizeof(double));
    c->rho2 = calloc(c->ly, sizeof(double));
    if (!c->L || !c->T || !c->kind || !c->p0 || !c->p1 || !c->prho || !c->rho1 || !c->rho2) {
        chan_free(c);
        return -1;
    }
    int fam_vel = cfg->bc_family ? K_ZH_VEL : K_REG_VEL;
    int fam_pres = cfg->bc_family ? K_ZH_PRES : K_REG_PRES;
    for (int y = 2; y <= c->ly - 1; ++y) c->prho[y] = 1.;
    int ox = cfg->obst_x, oy = cfg->obst_y, r = cfg->obst_r;
#pragma omp parallel for schedule(static)
    for (int x = 1; x <= c->lx; ++x)
        for (int y = 1; y <= c->ly; ++y) {
            double u = orc_poiseuille(y, c->ly, cfg->u_max);
            double usqr = u * u + 0. * 0.;
            double *f = node(c->L, c, x, y);
            for (int i = 0; i < 9; ++i) f[i] = feq(i, 1., u, 0., usqr);  /* lb.c:71-78 */
            size_t n = (size_t)x * c->ny + y;
            int inside = (x - ox) * (x - ox) + (y - oy) * (y - oy) <= r * r;
            c->kind[n] = inside ? K_BOUNCE : K_BULK;
        }
    for (int x = 1; x <= c->lx; ++x) {
        size_t lo = (size_t)x * c->ny + 1, up = (size_t)x * c->ny + c->ly;
        c->kind[lo] = fam_vel + LOWER; c->p0[lo] = 0.; c->p1[lo] = 0.;
        c->kind[up] = fam_vel + UPPER; c->p0[up] = 0.; c->p1[up] = 0.;
    }
    for (int y = 2; y <= c->ly - 1; ++y) {
        size_t le = (size_t)1 * c->ny + y, ri = (size_t)c->lx * c->ny + y;
        c->kind[le] = fam_vel + LEFT;
        c->p0[le] = orc_poiseuille(y, c->ly, cfg->u_max); c->p1[le] = 0.;
        c->kind[ri] = fam_pres + RIGHT;
        c->p0[ri] = c->prho[y]; c->p1[ri] = 0.;
    }
    return 0;
}

/* updateZeroGradientBoundary, unsteady.c:205-214: outlet density from a
 * second-order one-sided extrapolation of the current post-stream state. */
static void zgb_update(chan *c) {
    double r1, r2, u, v;
    for (int y = 2; y <= c->ly - 1; ++y) {
        macros(node(c->L, c, c->lx - 1, y), &r1, &u, &v);
        macros(node(c->L, c, c->lx - 2, y), &r2, &u, &v);
        c->prho[y] = 4. / 3. * r1 - 1. / 3. * r2;
    }
}

/* origin: collide everything in place, then stream into T and swap
 * (lb.c:172-201, 246-287).  combine (lb.c:353-390) performs the same
 * per-node work fused; its result is identical, so it shares this path. */
/* Both one-step orders run their column loop on OpenMP threads: a node's
 * collision reads and writes that node alone, and every slot a push writes
 * has exactly one source node, so the result does not depend on the thread
 * count or schedule (bit for bit). */
static void step_origin(chan *c) {
#pragma omp parallel for schedule(static)
    for (int x = 1; x <= c->lx; ++x)
        for (int y = 1; y <= c->ly; ++y) apply(c, c->L, x, y);
#pragma omp parallel for schedule(static)
    for (int x = 1; x <= c->lx; ++x)
        for (int y = 1; y <= c->ly; ++y) push(c, c->L, c->T, x, y);
    double *s = c->L; c->L = c->T; c->T = s;
}

static void step_combine(chan *c) {
#pragma omp parallel for schedule(static)
    for (int x = 1; x <= c->lx; ++x)
        for (int y = 1; y <= c->ly; ++y) {
            apply(c, c->L, x, y);
            push(c, c->L, c->T, x, y);
        }
    double *s = c->L; c->L = c->T; c->T = s;
}

/* Two steps per sweep (lb.c:442-572 and the blocked forms): the first
 * collide-and-stream of node (x, y) goes L -> T, and as soon as node
 * (x-1, y-1) has all its inputs in T its second collide-and-stream goes
 * T -> L.  Densities of columns lx-2 and lx-1 for the second step's outlet
 * are taken from T just before those nodes collide (lb.c:471-480). */
static void twostep_visit(chan *c, int x, int y) {
    apply(c, c->L, x, y);
    push(c, c->L, c->T, x, y);
    if (x > 1 && y > 1) {
        double u, v;
        if (x == c->lx - 1) macros(node(c->T, c, x - 1, y - 1), &c->rho2[y - 1], &u, &v);
        if (x == c->lx) macros(node(c->T, c, x - 1, y - 1), &c->rho1[y - 1], &u, &v);
        apply(c, c->T, x - 1, y - 1);
        push(c, c->T, c->L, x - 1, y - 1);
    }
}

/* the nodes every two-step order leaves for last: row ly, then column lx
 * with the extrapolated outlet density (lb.c:502-551) */
static void twostep_tail(chan *c) {
    int lx = c->lx, ly = c->ly;
    for (int x = 1; x < lx; ++x) { apply(c, c->T, x, ly); push(c, c->T, c->L, x, ly); }
    apply(c, c->T, lx, 1); push(c, c->T, c->L, lx, 1);
    for (int y = 2; y < ly; ++y) {
        if (c->zgb) c->prho[y] = 4. / 3. * c->rho1[y] - 1. / 3. * c->rho2[y];
        apply(c, c->T, lx, y); push(c, c->T, c->L, lx, y);
    }
    apply(c, c->T, lx, ly); push(c, c->T, c->L, lx, ly);
}

static void sweep_tight(chan *c) {                    /* lb.c:454-497 */
    for (int x = 1; x <= c->lx; ++x)
        for (int y = 1; y <= c->ly; ++y) twostep_visit(c, x, y);
    twostep_tail(c);
}

static void sweep_block(chan *c, int b) {             /* lb.c:950-996 */
    for (int X = 1; X <= c->lx; X += b)
        for (int Y = 1; Y <= c->ly; Y += b)
            for (int i = 0; i < b; ++i)
                for (int j = 0; j < b; ++j)
                    if (X + i <= c->lx && Y + j <= c->ly) twostep_visit(c, X + i, Y + j);
    twostep_tail(c);
}

static void sweep_panel(chan *c, int b) {             /* lb.c:1416-1465 */
    for (int Y = 1; Y <= c->ly; Y += b)
        for (int x = 1; x <= c->lx; ++x)
            for (int j = 0; j < b; ++j)
                if (Y + j <= c->ly) twostep_visit(c, x, Y + j);
    twostep_tail(c);
}

static void copy_out(const chan *c, const double *buf, double *out) {
    if (out) memcpy(out, buf, sizeof(double) * 9 * (size_t)(c->lx + 2) * c->ny);
}

static int run_chan(chan *c, const orc_cfg *cfg, double *lat_out, double *tmp_out,
                    double *prho_out) {
    int ran = 0;
    int twostep = cfg->variant >= 2;
    int b = cfg->blk > 0 ? cfg->blk : 1;
    /* unsteady.c:425-531 */
    for (int t = 0; t < cfg->steps; t += twostep ? 2 : 1) {
        if (c->zgb) zgb_update(c);
        switch (cfg->variant) {
        case 0: step_origin(c); break;
        case 1: step_combine(c); break;
        case 2: sweep_tight(c); break;
        case 3: sweep_block(c, b); break;
        default: sweep_panel(c, b); break;
        }
        ran += twostep ? 2 : 1;
    }
    copy_out(c, c->L, lat_out);
    copy_out(c, c->T, tmp_out);
    if (prho_out) memcpy(prho_out, c->prho, sizeof(double) * (c->ly + 2));
    return ran;
}

int orc_run(const orc_cfg *cfg, double *lat_out, double *tmp_out, double *prho_out) {
    chan c;
    if (cfg->lx < 3 || cfg->ly < 3 || cfg->variant < 0 || cfg->variant > 4) return -1;
    if (chan_init(&c, cfg)) return -2;
    int ran = run_chan(&c, cfg, lat_out, tmp_out, prho_out);
    chan_free(&c);
    return ran;
}

int orc_run_from(const orc_cfg *cfg, const double *lat_in, const double *tmp_in,
                 const double *prho_in, double *lat_out, double *tmp_out,
                 double *prho_out) {
    chan c;
    if (cfg->lx < 3 || cfg->ly < 3 || cfg->variant < 0 || cfg->variant > 4) return -1;
    if (chan_init(&c, cfg)) return -2;
    size_t nb = sizeof(double) * 9 * (size_t)(c.lx + 2) * c.ny;
    memcpy(c.L, lat_in, nb);
    memcpy(c.T, tmp_in, nb);
    if (prho_in) memcpy(c.prho, prho_in, sizeof(double) * (c.ly + 2));
    int ran = run_chan(&c, cfg, lat_out, tmp_out, prho_out);
    chan_free(&c);
    return ran;
}

/* ------------------------------------------------------------------ */
/* per-function entry points for known-answer tests                    */
/* ------------------------------------------------------------------ */

void orc_macros(const double *f, double *out3) { macros(f, &out3[0], &out3[1], &out3[2]); }

double orc_equilibrium(int i, double rho, double ux, double uy, double usqr) {
    return feq(i, rho, ux, uy, usqr);
}

void orc_bgk(double *f, double omega) { bgk_relax(f, omega); }

int orc_apply_bc(int which, double *f, double p0, double p1, double omega) {
    static const int map[17] = {
        K_BOUNCE,
        K_ZH_VEL + UPPER, K_ZH_VEL + LOWER, K_ZH_VEL + LEFT, K_ZH_VEL + RIGHT,
        K_ZH_PRES + UPPER, K_ZH_PRES + LOWER, K_ZH_PRES + LEFT, K_ZH_PRES + RIGHT,
        K_REG_VEL + UPPER, K_REG_VEL + LOWER, K_REG_VEL + LEFT, K_REG_VEL + RIGHT,
        K_REG_PRES + UPPER, K_REG_PRES + LOWER, K_REG_PRES + LEFT, K_REG_PRES + RIGHT};
    if (which < 0 || which > 16) return -1;
    collide_node(map[which], f, p0, p1, omega);
    return 0;
}

/* ------------------------------------------------------------------ */
/* diagnostics (no counterpart in the reference)                       */
/* ------------------------------------------------------------------ */

/* Momentum exchange on a post-stream state: the momentum carried into solid
 * nodes by populations that arrived from non-solid nodes, minus the momentum
 * carried out of solid nodes into non-solid nodes by populations that arrived
 * there.  Kinetic energy and mass are summed over non-solid interior nodes. */
void orc_diag(const orc_cfg *cfg, const double *lat, double *out4) {
    int lx = cfg->lx, ly = cfg->ly, ny = ly + 2;
    int ox = cfg->obst_x, oy = cfg->obst_y, r = cfg->obst_r;
#define SOLID(x, y) ((x) >= 1 && (x) <= lx && (y) >= 2 && (y) <= ly - 1 && \
                     ((x) - ox) * ((x) - ox) + ((y) - oy) * ((y) - oy) <= r * r && \
                     (x) != 1 && (x) != lx)
    double fx = 0., fy = 0., ke = 0., mass = 0.;
    for (int x = 1; x <= lx; ++x)
        for (int y = 1; y <= ly; ++y) {
            const double *f = lat + ((size_t)x * ny + y) * 9;
            int s = SOLID(x, y);
            if (!s) {
                double rho, ux, uy;
                macros(f, &rho, &ux, &uy);
                ke += 0.5 * rho * (ux * ux + uy * uy);
                mass += rho;
            }
            for (int i = 1; i < 9; ++i) {
                int sx = x - CX[i], sy = y - CY[i];
                if (sx < 1 || sx > lx || sy < 1 || sy > ly) continue;
                int from_solid = SOLID(sx, sy);
                if (s && !from_solid) { fx += CX[i] * f[i]; fy += CY[i] * f[i]; }
                if (!s && from_solid) { fx -= CX[i] * f[i]; fy -= CY[i] * f[i]; }
            }
        }
#undef SOLID
    out4[0] = fx; out4[1] = fy; out4[2] = ke; out4[3] = mass;
}

/* Periodic ghost copy on a post-stream state (lb.c:1783-1817): what left the
 * domain through one side, and was pushed into that side's ghost ring,
 * re-enters at the opposite side.  Row pass, column pass, then the four
 * corner diagonals, in the reference's order (a corner slot written by a
 * pass is overwritten by its diagonal). */
void orc_make_periodic(int lx, int ly, double *lat) {
    const int ny = ly + 2;
#define POP(x, y, i) lat[((size_t)(x) * ny + (y)) * 9 + (i)]
    for (int x = 1; x <= lx; ++x) {
        POP(x, ly, 4) = POP(x, 0, 4);          /* leaving downwards re-enters at the top */
        POP(x, ly, 7) = POP(x, 0, 7);
        POP(x, ly, 8) = POP(x, 0, 8);
        POP(x, 1, 2) = POP(x, ly + 1, 2);      /* leaving upwards re-enters at the bottom */
        POP(x, 1, 5) = POP(x, ly + 1, 5);
        POP(x, 1, 6) = POP(x, ly + 1, 6);
    }
    for (int y = 1; y <= ly; ++y) {
        POP(1, y, 1) = POP(lx + 1, y, 1);      /* leaving right re-enters on the left */
        POP(1, y, 5) = POP(lx + 1, y, 5);
        POP(1, y, 8) = POP(lx + 1, y, 8);
        POP(lx, y, 3) = POP(0, y, 3);          /* leaving left re-enters on the right */
        POP(lx, y, 6) = POP(0, y, 6);
        POP(lx, y, 7) = POP(0, y, 7);
    }
    POP(1, 1, 5) = POP(lx + 1, ly + 1, 5);
    POP(lx, 1, 6) = POP(0, ly + 1, 6);
    POP(lx, ly, 7) = POP(0, 0, 7);
    POP(1, ly, 8) = POP(lx + 1, 0, 8);
#undef POP
}
