/*
 * boundaries.h -- straight-wall node operators, drop-in for the reference's
 * boundaries.h (IUPU-HPC/memory-aware-LBM boundaries.h:42-102).
 *
 * The parameter structs keep the reference layouts; the 17 operators keep
 * their names and signatures.  Their addresses identify the dynamics of a
 * node when lbm_bind_device translates a Simulation for the GPU.  Called
 * directly on a host population array they run the same device operator on
 * that one node (through lbm_hip_apply_node_op), so host and device agree bit
 * for bit.
 */
#ifndef BOUNDARIES_H
#define BOUNDARIES_H

#include "lb.h"

/* Dirichlet velocity value and the bulk dynamics (boundaries.h:50-53) */
typedef struct {
    Dynamics *bulkDynamics;
    double ux, uy;
} VelocityBCData;

/* prescribed density, tangential velocity, bulk dynamics (boundaries.h:64-67) */
typedef struct {
    Dynamics *bulkDynamics;
    double rho, uPar;
} PressureBCData;

void bounceBack(double *fPop, void *selfData);

void upperZouHe(double *fPop, void *selfData);
void lowerZouHe(double *fPop, void *selfData);
void leftZouHe(double *fPop, void *selfData);
void rightZouHe(double *fPop, void *selfData);

void upperPressureZouHe(double *fPop, void *selfData);
void lowerPressureZouHe(double *fPop, void *selfData);
void leftPressureZouHe(double *fPop, void *selfData);
void rightPressureZouHe(double *fPop, void *selfData);

void upperRegularized(double *fPop, void *selfData);
void lowerRegularized(double *fPop, void *selfData);
void leftRegularized(double *fPop, void *selfData);
void rightRegularized(double *fPop, void *selfData);

void upperPressureRegularized(double *fPop, void *selfData);
void lowerPressureRegularized(double *fPop, void *selfData);
void leftPressureRegularized(double *fPop, void *selfData);
void rightPressureRegularized(double *fPop, void *selfData);

#endif /* BOUNDARIES_H */
