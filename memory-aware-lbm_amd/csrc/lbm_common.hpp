// lbm_common.hpp -- device data layout and node classification shared by
// every kernel of the library.
//
// Layout in HBM (structure of arrays with the populations of a column side
// by side):
//   element (i, x, y) at  i*plane + x*pitch + y + yoff,  pitch = 9*plane
// with x = 0..lx+1 and y = 0..ly+1 (a one-node ghost ring, as the
// reference's (lx+2)*(ly+2) buffers, lb.c:89-100).  y is the contiguous
// dimension, like the reference's lattice[iX][iY]; `plane` is one
// population's segment of a column (ly + 2 + yoff elements rounded up to
// ALIGN), and the nine segments of column x follow each other.  yoff =
// ALIGN-1 places y = 1 on an ALIGN-element boundary, so every segment's
// interior starts on a 512-byte (FP64) boundary.  A sweep that reads or
// writes the nine populations of a column then touches one contiguous
// 9 * plane-element region instead of nine regions a whole population apart:
// on 16384 x 4096 the fused FP64 step runs 3.8% faster, the two-step tiles
// 15%, the register marches 1-3%, and a plain nine-stream copy of the same
// shape moves 5.70 TB/s against 5.39 for nine separate planes (MI355X).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

constexpr int ALIGN = 64;  // elements

struct Layout {
    int lx, ly;
    int pitch;        // elements from column x to column x+1 (all nine populations)
    int yoff;         // offset of y = 0 inside a population's segment
    int64_t plane;    // elements from population i to i+1 of the same node

    __host__ __device__ int64_t at(int i, int x, int y) const {
        return (int64_t)i * plane + (int64_t)x * pitch + y + yoff;
    }
    __host__ __device__ int64_t node(int x, int y) const {
        return (int64_t)x * pitch + y + yoff;
    }
    // elements of one whole buffer (nine populations, ghost ring included)
    __host__ __device__ int64_t span() const { return (int64_t)(lx + 2) * pitch; }
    // contiguous elements of one population in one column
    __host__ __device__ int64_t segment() const { return plane < pitch ? plane : pitch; }
    // the node's index in the reference's (lx+2) x (ly+2) arrays: a key that
    // does not depend on the device layout (FP32 stochastic rounding)
    __host__ __device__ int64_t key(int x, int y) const { return (int64_t)x * (ly + 2) + y; }
};

inline Layout make_layout(int lx, int ly) {
    Layout L;
    L.lx = lx;
    L.ly = ly;
    L.yoff = ALIGN - 1;
    const int need = L.yoff + ly + 2;
    const int seg = (need + ALIGN - 1) / ALIGN * ALIGN;
    L.plane = seg;
    L.pitch = 9 * seg;
    return L;
}

// Node classes of the cylinder channel, resolved analytically from (x, y) so
// the hot loop reads no class map (the reference dispatches through a
// per-node Dynamics pointer instead, lb.c:125-128; its assignment order is
// unsteady.c:162-199: disc, then rows 1 and ly, then columns 1 and lx).
enum NodeClass : int { NC_BULK = 0, NC_SOLID = 1, NC_WALL = 2 };

struct SideBC {
    int kind;            // BcKind
    int family;          // BcFamily
    const double *p0;    // device, indexed by x (upper/lower) or y (left/right)
    const double *p1;
};

struct Problem {
    Layout lay;
    int ox, oy, r2;      // disc; r2 < 0 means no disc
    double omega;
    SideBC side[4];
    int zgb;             // right side density by extrapolation
    double *outlet_rho;  // device, ly+2: outlet density used by the last step
    // device counter of the collisions done so far (reduced storage only: it
    // keys the stochastic rounding); a launch reads it, the next one bumps it
    const uint32_t *step;
};

__host__ __device__ inline bool in_disc(const Problem &p, int x, int y) {
    int dx = x - p.ox, dy = y - p.oy;
    return dx * dx + dy * dy <= p.r2;
}

// side of a perimeter node, or -1 for interior nodes
__host__ __device__ inline int side_of(const Problem &p, int x, int y) {
    if (y == 1) return 1;              // LOWER
    if (y == p.lay.ly) return 0;       // UPPER
    if (x == 1) return 2;              // LEFT
    if (x == p.lay.lx) return 3;       // RIGHT
    return -1;
}

// Workgroups are dealt to the 8 XCDs in turn; renumber them so that each XCD
// takes a contiguous range of [0, nblk) (a bijection): neighbouring tiles or
// bands, whose halo rows and shared cache lines overlap, then meet in the
// same L2.
__device__ __forceinline__ int xcd_block(int b, int nblk) {
    const int xcd = b % 8, k = b / 8, per = nblk / 8, rem = nblk % 8;
    return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

}  // namespace lbm
