/*
 * boundaries.c -- host side of include/boundaries.h (the reference's
 * boundaries.h:72-102 operators).
 *
 * Each operator's address identifies a node's dynamics when a Simulation is
 * translated for the GPU.  Called directly on one host node, it runs the
 * corresponding device operator on that node, so a driver that post-processes
 * on the host gets the device's exact numbers; the node goes through the
 * library's fixed staging array (lbm_hip_apply_node_op_noalloc: no device
 * allocation and no device-wide wait per call).
 */
#include <stdio.h>
#include <stdlib.h>

#include "../../include/boundaries.h"
#include "../../include/lbm_hip.h"

/* boundaries.h order, matching lbm_hip_apply_node_op(_noalloc)'s `which` */
enum {
    OP_BOUNCE = 0,
    OP_ZH_VEL = 1,        /* + side: upper, lower, left, right */
    OP_ZH_PRES = 5,
    OP_REG_VEL = 9,
    OP_REG_PRES = 13
};
enum { S_UPPER = 0, S_LOWER = 1, S_LEFT = 2, S_RIGHT = 3 };

static double bulk_omega_of(const Dynamics *bulk, const char *fn) {
    if (!bulk || bulk->dynamicsFun != bgk || !bulk->selfData) {
        fprintf(stderr, "%s: the bulk dynamics must be bgk\n", fn);
        abort();
    }
    return *(const double *)bulk->selfData;
}

static void run(int which, double *f, double p0, double p1, double omega, const char *fn) {
    int rc = lbm_hip_apply_node_op_noalloc(which, 1, f, &p0, &p1, omega);
    if (rc) {
        fprintf(stderr, "%s: device error %d: %s\n", fn, rc, lbm_hip_last_error());
        abort();
    }
}

static void velocity_op(int which, double *f, void *self, const char *fn) {
    const VelocityBCData *d = self;
    run(which, f, d->ux, d->uy, bulk_omega_of(d->bulkDynamics, fn), fn);
}

static void pressure_op(int which, double *f, void *self, const char *fn) {
    const PressureBCData *d = self;
    run(which, f, d->rho, d->uPar, bulk_omega_of(d->bulkDynamics, fn), fn);
}

void bounceBack(double *fPop, void *selfData) {                 /* boundaries.c:48-59 */
    (void)selfData;
    run(OP_BOUNCE, fPop, 0., 0., 1., "bounceBack");
}

/* Zou/He velocity and pressure operators: boundaries.c:171-233 */
void upperZouHe(double *f, void *s) { velocity_op(OP_ZH_VEL + S_UPPER, f, s, "upperZouHe"); }
void lowerZouHe(double *f, void *s) { velocity_op(OP_ZH_VEL + S_LOWER, f, s, "lowerZouHe"); }
void leftZouHe(double *f, void *s) { velocity_op(OP_ZH_VEL + S_LEFT, f, s, "leftZouHe"); }
void rightZouHe(double *f, void *s) { velocity_op(OP_ZH_VEL + S_RIGHT, f, s, "rightZouHe"); }
void upperPressureZouHe(double *f, void *s) {
    pressure_op(OP_ZH_PRES + S_UPPER, f, s, "upperPressureZouHe");
}
void lowerPressureZouHe(double *f, void *s) {
    pressure_op(OP_ZH_PRES + S_LOWER, f, s, "lowerPressureZouHe");
}
void leftPressureZouHe(double *f, void *s) {
    pressure_op(OP_ZH_PRES + S_LEFT, f, s, "leftPressureZouHe");
}
void rightPressureZouHe(double *f, void *s) {
    pressure_op(OP_ZH_PRES + S_RIGHT, f, s, "rightPressureZouHe");
}

/* regularized velocity and pressure operators: boundaries.c:304-414 */
void upperRegularized(double *f, void *s) {
    velocity_op(OP_REG_VEL + S_UPPER, f, s, "upperRegularized");
}
void lowerRegularized(double *f, void *s) {
    velocity_op(OP_REG_VEL + S_LOWER, f, s, "lowerRegularized");
}
void leftRegularized(double *f, void *s) {
    velocity_op(OP_REG_VEL + S_LEFT, f, s, "leftRegularized");
}
void rightRegularized(double *f, void *s) {
    velocity_op(OP_REG_VEL + S_RIGHT, f, s, "rightRegularized");
}
void upperPressureRegularized(double *f, void *s) {
    pressure_op(OP_REG_PRES + S_UPPER, f, s, "upperPressureRegularized");
}
void lowerPressureRegularized(double *f, void *s) {
    pressure_op(OP_REG_PRES + S_LOWER, f, s, "lowerPressureRegularized");
}
void leftPressureRegularized(double *f, void *s) {
    pressure_op(OP_REG_PRES + S_LEFT, f, s, "leftPressureRegularized");
}
void rightPressureRegularized(double *f, void *s) {
    pressure_op(OP_REG_PRES + S_RIGHT, f, s, "rightPressureRegularized");
}
