/*
 * unsteady_gpu.c -- the cylinder-channel driver with the reference's command
 * line, running its sweeps on the GPU through include/lb.h.
 *
 *   STOP=<steps> ./unsteady_gpu lx ly blk_size [thread_block] [variant]
 *
 * mirrors `unsteady_<variant> lx ly blk_size thread_block` of the reference
 * (unsteady.c:84-109, 241-565): the same problem set-up (disc at (lx/4, ly/2)
 * with radius ly/10+1, Re 100, u_max 0.02, regularized walls, Poiseuille
 * inlet, pressure outlet with zero-gradient density), the same step loop and
 * the same progress and timing lines.  variant is one of origin, combine,
 * tight, tight_block, panel.  It is taken from argv[5] if
 * given, else from the program's own name when that is one of the
 * reference's executables, unsteady_<variant>[_openmp][_papi]
 * (CMakeLists.txt:27-50; the Makefile installs those names as links to this
 * program under bin/, so scripts such as run_exps.sh:67-76 run unchanged; the
 * _openmp and _papi suffixes select the same device sweep), else from
 * $LBM_VARIANT, else origin.  The _papi names print no counter lines: the
 * counter report of a sweep is `python -m lbm_amd.counters`.
 *
 * Extra environment (none needed for the reference behaviour):
 *   LBM_DISC=B        disc (lx/5, ly/2, ly/10) instead of the driver's own
 *   LBM_FAMILY=zouhe  Zou/He walls instead of regularized (unsteady.c:60-61)
 *   LBM_DUMP=<file>   after the run, write lattice then tmpLattice as raw
 *                     doubles in node layout ((lx+2)*(ly+2)*9 each)
 *   LBM_SAVEVEL=<file> saveVel() of the final state
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "../../include/boundaries.h"
#include "../../include/lb.h"
#include "../../include/lbm_hip.h"

int lx, ly;
int obst_x, obst_y, obst_r;
double uMax, Re, nu, omega;
int maxT;
int blk_size, thread_block;

static Dynamics bulkDynamics = {&bgk, (void *)&omega};
static Dynamics bounceBackDynamics = {&bounceBack, 0};
static VelocityBCData zeroVelocityBoundary = {&bulkDynamics, 0., 0.};
static Dynamics upperBoundary, lowerBoundary;
static Dynamics *leftBoundary, *rightBoundary;
static VelocityBCData *poiseuilleBoundary;
PressureBCData *pressureBoundary;
static Simulation sim;

static double wall_time(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + tv.tv_usec / 1000000.0;
}

/* the reference driver's outlet update (unsteady.c:205-214); with a bound
 * Simulation the device performs it inside every sweep instead */
void updateZeroGradientBoundary() {
    double r1, u1, v1, r2, u2, v2;
    for (int y = 2; y <= ly - 1; ++y) {
        computeMacros(sim.lattice[lx - 1][y].fPop, &r1, &u1, &v1);
        computeMacros(sim.lattice[lx - 2][y].fPop, &r2, &u2, &v2);
        pressureBoundary[y].rho = 4. / 3. * r1 - 1. / 3. * r2;
        pressureBoundary[y].uPar = 0.;
    }
}

static double poiseuille(int y) {                       /* unsteady.c:155-159 */
    double yy = (double)(y - 1);
    double L = (double)(ly - 1);
    return 4. * uMax / (L * L) * (L * yy - yy * yy);
}

/* the sweep named by a reference executable name, unsteady_<variant>
 * [_openmp][_papi] (CMakeLists.txt:27-50), or NULL */
static const char *variant_of_name(const char *path) {
    static const char *const names[] = {"origin", "combine", "tight_block", "tight", "panel"};
    static char v[32];
    const char *base = strrchr(path, '/');
    base = base ? base + 1 : path;
    if (strncmp(base, "unsteady_", 9)) return NULL;
    base += 9;
    for (size_t k = 0; k < sizeof names / sizeof names[0]; ++k) {
        size_t n = strlen(names[k]);
        const char *rest = base + n;
        if (strncmp(base, names[k], n)) continue;
        if (*rest && strcmp(rest, "_openmp") && strcmp(rest, "_papi") &&
            strcmp(rest, "_openmp_papi"))
            continue;
        snprintf(v, sizeof v, "%s", names[k]);
        return v;
    }
    return NULL;
}

int main(int argc, char *argv[]) {
    if (argc < 4) {
        fprintf(stderr, "usage: STOP=<steps> %s lx ly blk_size [thread_block] [variant]\n",
                argv[0]);
        return 2;
    }
    const char *stop = getenv("STOP");
    if (!stop) { fprintf(stderr, "STOP is not set\n"); return 2; }
    lx = atoi(argv[1]);
    ly = atoi(argv[2]);
    blk_size = atoi(argv[3]);
    thread_block = argc > 4 ? atoi(argv[4]) : 0;
    const char *variant = argc > 5 ? argv[5] : variant_of_name(argv[0]);
    if (!variant) variant = getenv("LBM_VARIANT");
    if (!variant) variant = "origin";
    const char *disc = getenv("LBM_DISC");
    int zouhe = getenv("LBM_FAMILY") && !strcmp(getenv("LBM_FAMILY"), "zouhe");

    /* unsteady.c:84-109 */
    if (disc && !strcmp(disc, "B")) {
        obst_r = ly / 10; obst_x = lx / 5; obst_y = ly / 2;
    } else {
        obst_r = ly / 10 + 1; obst_x = lx / 4; obst_y = ly / 2;
    }
    uMax = 0.02;
    Re = 100;
    nu = uMax * 2. * obst_r / Re;
    omega = 1. / (3 * nu + 1. / 2.);
    maxT = atoi(stop);
    int interval = maxT / 10 > 0 ? maxT / 10 : 1;
    printf("\nlx=%d, ly=%d, omega=%f, blk_size=%d\n\n", lx, ly, omega, blk_size);

    int twosteps = 0;
    void (*collision_func)(Simulation *) = NULL, (*stream_func)(Simulation *) = NULL;
    if (!strcmp(variant, "combine")) {
        collision_func = collide_with_stream; stream_func = finalize_stream;
    } else if (!strcmp(variant, "tight")) {
        collision_func = collide_tight; twosteps = 1;
    } else if (!strcmp(variant, "tight_block")) {
        collision_func = collide_tight_block; twosteps = 1;
    } else if (!strcmp(variant, "panel")) {
        collision_func = collide_tight_panel_iy; twosteps = 1;
    } else if (!strcmp(variant, "origin")) {
        collision_func = collide; stream_func = propagate;
    } else {
        fprintf(stderr, "unknown variant %s\n", variant);
        return 2;
    }
    printf("use case %s (MI355X)\n", variant);
    fflush(stdout);

    /* unsteady.c:113-142 */
    upperBoundary.dynamicsFun = zouhe ? upperZouHe : upperRegularized;
    upperBoundary.selfData = &zeroVelocityBoundary;
    lowerBoundary.dynamicsFun = zouhe ? lowerZouHe : lowerRegularized;
    lowerBoundary.selfData = &zeroVelocityBoundary;
    poiseuilleBoundary = calloc(ly + 2, sizeof(VelocityBCData));
    pressureBoundary = calloc(ly + 2, sizeof(PressureBCData));
    leftBoundary = calloc(ly + 2, sizeof(Dynamics));
    rightBoundary = calloc(ly + 2, sizeof(Dynamics));
    for (int y = 2; y <= ly - 1; ++y) {
        poiseuilleBoundary[y].bulkDynamics = &bulkDynamics;
        poiseuilleBoundary[y].uy = 0.;
        pressureBoundary[y].bulkDynamics = &bulkDynamics;
        pressureBoundary[y].rho = 1.;
        pressureBoundary[y].uPar = 0.;
        leftBoundary[y].dynamicsFun = zouhe ? leftZouHe : leftRegularized;
        leftBoundary[y].selfData = &poiseuilleBoundary[y];
        rightBoundary[y].dynamicsFun = zouhe ? rightPressureZouHe : rightPressureRegularized;
        rightBoundary[y].selfData = &pressureBoundary[y];
    }

    constructSim(&sim, lx, ly);
    /* unsteady.c:162-199 */
    for (int x = 1; x <= lx; ++x)
        for (int y = 1; y <= ly; ++y) {
            iniEquilibrium(&sim.lattice[x][y], 1., poiseuille(y), 0.);
            if ((x - obst_x) * (x - obst_x) + (y - obst_y) * (y - obst_y) <= obst_r * obst_r)
                setDynamics(&sim, x, y, &bounceBackDynamics);
            else
                setDynamics(&sim, x, y, &bulkDynamics);
        }
    for (int x = 1; x <= lx; ++x) {
        setDynamics(&sim, x, 1, &lowerBoundary);
        setDynamics(&sim, x, ly, &upperBoundary);
    }
    for (int y = 2; y <= ly - 1; ++y) {
        poiseuilleBoundary[y].ux = poiseuille(y);
        setDynamics(&sim, 1, y, &leftBoundary[y]);
        setDynamics(&sim, lx, y, &rightBoundary[y]);
    }
    printf("Pass iniGeometry, obst_x=%d, obst_y=%d, obst_r=%d\n", obst_x, obst_y, obst_r);
    fflush(stdout);

    lbm_bind_opts opts = {LBM_STORE_F64, -1, 0, 0};
    if (lbm_bind_device_ex(&sim, &opts)) {
        fprintf(stderr, "lbm_bind_device: %s\n", lbm_last_error());
        return 1;
    }

    double t_start = wall_time(), t_collid = 0., t_stream = 0.;
    int ran = 0;
    for (int iT = 0; iT < maxT; iT += twosteps ? 2 : 1) {   /* unsteady.c:425-531 */
        if (iT % interval == 0) { printf("t=%d\n", iT); fflush(stdout); }
        double t0 = wall_time();
        collision_func(&sim);
        double t1 = wall_time();
        t_collid += t1 - t0;
        if (!twosteps) {
            t0 = wall_time();
            stream_func(&sim);
            t_stream += wall_time() - t0;
        }
        ran += twosteps ? 2 : 1;
    }
    double t0 = wall_time();
    lbm_hip_sync(lbm_device_of(&sim));
    t_collid += wall_time() - t0;
    double t_end = wall_time();
    printf("time eclapsed in %d steps, t_total=%.3f, t_collid=%.3f, t_stream=%.3f, t_zgb=%.3f\n",
           maxT, t_end - t_start, t_collid, t_stream, 0.0);
    printf("MLUPS=%.1f (%d steps run)\n", (double)lx * ly * ran / (t_end - t_start) / 1e6, ran);

    const char *dump = getenv("LBM_DUMP");
    if (dump) {
        if (lbm_sync_host(&sim)) { fprintf(stderr, "sync: %s\n", lbm_last_error()); return 1; }
        FILE *o = fopen(dump, "wb");
        if (!o) { perror(dump); return 1; }
        Node **bufs[2] = {sim.lattice, sim.tmpLattice};
        for (int b = 0; b < 2; ++b)
            for (int x = 0; x < lx + 2; ++x)
                for (int y = 0; y < ly + 2; ++y) fwrite(bufs[b][x][y].fPop, sizeof(double), 9, o);
        fclose(o);
    }
    const char *sv = getenv("LBM_SAVEVEL");
    if (sv) saveVel(&sim, (char *)sv);

    destructSim(&sim);
    free(poiseuilleBoundary);
    free(pressureBoundary);
    free(leftBoundary);
    free(rightBoundary);
    return 0;
}
