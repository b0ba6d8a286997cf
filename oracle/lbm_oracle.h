/*
 * lbm_oracle.h -- CPU oracle for the D2Q9 BGK collide-stream path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product library never links it and has no CPU fallback to it.
 *
 * A clean-room restatement of the reference's algorithm
 * (IUPU-HPC/memory-aware-LBM lb.c, boundaries.c, unsteady.c); every function
 * cites the reference lines it follows.  Floating-point expressions keep the
 * reference's evaluation order so that results are bit-identical when built
 * with -ffp-contract=off.
 *
 * Pinning (what fixes each function independently of this text):
 *   orc_macros, orc_equilibrium, orc_bgk, orc_apply_bc (all 17 operators):
 *       bitwise known-answer vectors written by the real reference
 *       (tests/golden/kat.npz), plus moment closed forms and invariants.
 *   orc_run (all five sweep orders, both BC families, ZGB, stale slots):
 *       bitwise golden lattices written by the real reference
 *       (tests/golden/traj_*.npz, c1_*.npz), plus the quiescent-channel and
 *       variant-equivalence properties.
 *   orc_make_periodic: bitwise against the reference's makePeriodic on a
 *       seeded random lattice and inside an origin trajectory
 *       (tests/golden/periodic.npz).
 *   orc_diag (drag/lift by momentum exchange, kinetic energy): the reference
 *       has no such routine ("parity unpinned" against the reference); it is
 *       pinned by closed forms in tests/test_oracle.py instead.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int lx, ly;
    int obst_x, obst_y, obst_r;
    double u_max, reynolds;
    int bc_family;  /* 0 regularized, 1 Zou/He */
    int variant;    /* 0 origin, 1 combine, 2 tight, 3 tight_block, 4 panel */
    int openmp;     /* ignored: the oracle is serial */
    int nthreads;   /* ignored */
    int blk;        /* block / panel size for variants 3 and 4 */
    int tb;         /* ignored */
    int steps;
    int zgb;        /* 1: zero-gradient outlet density every step */
} orc_cfg;

/* Relaxation parameter of a configuration: unsteady.c:102-103 */
double orc_omega_of(double u_max, int obst_r, double reynolds);
/* Poiseuille inlet profile value: unsteady.c:155-159 */
double orc_poiseuille(int y, int ly, double u_max);

/* Run a full simulation; arrays are (lx+2)*(ly+2)*9 doubles, x slowest,
 * population fastest, ghost ring included (the reference Node layout minus
 * the dynamics pointer).  Returns the number of steps run. */
int orc_run(const orc_cfg *cfg, double *lat_out, double *tmp_out, double *prho_out);

/* Same as orc_run but starting from given buffers (lattice, tmpLattice) and a
 * given outlet density array; used to resume from a downloaded state. */
int orc_run_from(const orc_cfg *cfg, const double *lat_in, const double *tmp_in,
                 const double *prho_in, double *lat_out, double *tmp_out,
                 double *prho_out);

void orc_macros(const double *f, double *out3);
double orc_equilibrium(int i, double rho, double ux, double uy, double usqr);
void orc_bgk(double *f, double omega);
/* which: 0 bounce-back; 1..4 Zou/He velocity upper/lower/left/right;
 * 5..8 Zou/He pressure u/l/l/r; 9..12 regularized velocity; 13..16
 * regularized pressure. */
int orc_apply_bc(int which, double *f, double p0, double p1, double omega);

/* Diagnostics on a post-stream lattice (same layout as orc_run's output):
 * out[0] = drag (x force on the disc by momentum exchange),
 * out[1] = lift (y force), out[2] = kinetic energy sum(rho*|u|^2)/2 over
 * fluid nodes, out[3] = total mass over fluid nodes. */
void orc_diag(const orc_cfg *cfg, const double *lat, double *out4);

/* Periodic ghost copy (lb.c:1783-1817) on a lattice of the orc_run layout,
 * in place. */
void orc_make_periodic(int lx, int ly, double *lat);

#ifdef __cplusplus
}
#endif
#endif
