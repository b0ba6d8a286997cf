// launch.hpp -- host-side launchers of the device kernels (C++ linkage,
// private to the library; the public surface is include/lbm_hip.h).
#pragma once

#include <hip/hip_runtime.h>
#include "lbm_common.hpp"

namespace lbm {

// Optional timing of the kernel that carries a launch's lattice updates
// (measurement support).  A launcher given a probe records t0 and t1 on that
// kernel's own stream right before and after launching it, and fills in the
// kernel's name (as the profiler prints it) and the nodes it updates per
// launch.
struct KernelProbe {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    char name[64] = {0};
    double nodes = 0.;
};

// FP64 (bit-exact) path: kernels_f64.hip
hipError_t launch_collide_f64(const Problem &p, bool pull, const double *src, double *dst,
                              hipStream_t st);
hipError_t launch_stream_f64(const Layout &L, const double *src, double *dst, hipStream_t st);
// rows per lane of the one-step kernels on this lattice (storage 0 / 1 / 2:
// FP64 / FP32 / FP16), the last template argument of their names
int one_step_rows(const Layout &L, int storage);
hipError_t launch_init_f64(const Layout &L, const double *init_ux, double *dst, hipStream_t st);
// moments of columns x0 .. x0 + nx - 1 (x-major output from column x0)
hipError_t launch_macros_f64(const Layout &L, int x0, int nx, bool pull, const double *src,
                             double *rho, double *ux, double *uy, hipStream_t st);
hipError_t launch_copy(const void *src, void *dst, int64_t bytes, hipStream_t st);
// periodic ghost copy on a streamed state (L from T), any storage format
hipError_t launch_periodic(int storage, const Layout &L, const void *T, void *Lb,
                           hipStream_t st);
hipError_t launch_bump(uint32_t *ctr, uint32_t by, hipStream_t st);
int64_t diag_partials_needed(const Layout &L);
hipError_t launch_diag_f64(const Problem &p, bool pull, const double *src, double *partial,
                           double *out, hipStream_t st);
hipError_t launch_node_op_f64(int which, int n, double *f, const double *p0, const double *p1,
                              double omega, hipStream_t st);

hipError_t launch_diag_final(const double *partial, int64_t nblocks, double *out, hipStream_t st);

// Reduced-precision storage (FP32 = 1, FP16 = 2): kernels_lowp.hip
hipError_t launch_collide_lp(int storage, const Problem &p, bool pull, const void *src, void *dst,
                             hipStream_t st);
hipError_t launch_stream_lp(int storage, const Layout &L, const void *src, void *dst,
                            hipStream_t st);
hipError_t launch_init_lp(int storage, const Layout &L, const double *init_ux, void *dst,
                          hipStream_t st);
hipError_t launch_macros_lp(int storage, const Layout &L, int x0, int nx, bool pull,
                            const void *src, double *rho, double *ux, double *uy, hipStream_t st);
hipError_t launch_diag_lp(int storage, const Problem &p, bool pull, const void *src,
                          double *partial, hipStream_t st);
hipError_t launch_fill_zero_f(int storage, const Layout &L, void *dst, hipStream_t st);
hipError_t launch_widen(int storage, const Layout &L, const void *src, double *dst, hipStream_t st);
hipError_t launch_narrow(int storage, const Layout &L, const double *src, void *dst, hipStream_t st);

// Two-step (temporally blocked) kernels: kernels_2step_f64.hip
struct BlockGrid {
    int tx;      // tile columns (>= 3)
    int typ;     // tile rows + 2 (32, 64 or 128)
    int tbytes;  // bytes per LDS value (8: FP64 storage, 4: FP32 / FP16 storage)
    int nbx, nby;
};
BlockGrid block_grid(const Layout &L, int tx, int typ, int tbytes);
size_t block_lds_bytes(const BlockGrid &g);
hipError_t launch_block(int storage, const Problem &p, const BlockGrid &g, const void *P, void *Q,
                        hipStream_t st);
hipError_t block_kernel_setup();

// Two steps per launch, first step held in wave64 registers: kernels_wave.hip
constexpr int WAVE_ROWS = 62;   // owned rows per wave (64 lanes incl. a halo row each side)
struct WaveGrid {
    int cx;                        // columns per chunk (>= 3)
    int nbands, ncx, ntasks;
    int b_lo, b_hi, c_lo, c_hi;    // (band, chunk) rectangle of the interior waves
    int nin, nedge;                // interior waves, perimeter waves
    int r0, r1, nbl;               // interior rows [r0, r1) in nbl bands of the
                                   // interior kernel's height
    // thin frame: the interior rows reach as close to the walls as their
    // halo allows, the frame keeps the few rows below r0 and from r1 up
    // (strips of at most six rows, gl lanes per chunk, 64 / gl chunks per
    // perimeter wave: nthin waves) and the edge chunks of rows [r0, r1) in
    // 62-row bands
    int thin, nthin, gl;
    // interior chunks are right-aligned at column xr (exclusive end of the
    // last chunk; lx + 1 for the single-lattice grids); a chunk that would
    // start left of xl is moved right to start at xl (it then overlaps its
    // neighbour, which writes the same values)
    int xr, xl;
};

// right end (exclusive) of interior chunk ck of g
__host__ __device__ inline int in_chunk_x1(const WaveGrid &g, int ck) {
    const int X1 = g.xr - (g.ncx - 1 - ck) * g.cx;
    return X1 - g.cx < g.xl ? g.xl + g.cx : X1;
}

// Rows [Y0-1, Y0+62] (lanes) and columns [X0, X1) owned by wave (band, ck);
// chunks are aligned to the right end of the channel.
struct WaveSpan { int Y0, X0, X1; };
__host__ __device__ inline WaveSpan wave_span(const Layout &L, const WaveGrid &g, int band, int ck) {
    WaveSpan w;
    w.Y0 = 1 + band * WAVE_ROWS;
    w.X1 = L.lx + 1 - (g.ncx - 1 - ck) * g.cx;
    w.X0 = w.X1 - g.cx > 1 ? w.X1 - g.cx : 1;
    return w;
}

WaveGrid wave_grid(const Layout &L, int cx);
// false for lattices so tall that even three-column chunks overflow the
// interior kernel's 32-bit offsets (ly above about 1.9 M rows): the register
// marches refuse them
bool wave_grid_fits(const Layout &L);
hipError_t launch_wpanel(int storage, const Problem &p, const WaveGrid &g, const void *P, void *Q,
                         hipStream_t st, hipStream_t side, hipEvent_t fork, hipEvent_t join,
                         KernelProbe *probe = nullptr);
// LEVELS = 4, 6 or 8 steps per launch: the interior waves march LEVELS levels
// in registers (P -> Q); the perimeter frame (DeepFrame) runs LEVELS/2
// two-step launches through LEVELS/2 - 1 further buffers mid[] (P -> mid[0]
// -> ... -> Q).  Returns hipErrorInvalidValue when the lattice has no
// interior for it (wpanel_deep_possible: an interior of at least 124 rows
// and two chunks of the thin-frame grid).
bool wpanel_deep_possible(const Layout &L, int cx, int storage, int levels);
// the interior chunk width a deep launch takes (cx < 3: chosen for the lattice)
int wpanel_deep_cx(const Layout &L, int cx, int storage, int levels);
hipError_t launch_wpanel_deep(int storage, const Problem &p, int cx, int levels, const void *P,
                              void *const *mid, void *Q, hipStream_t st, hipStream_t side,
                              hipEvent_t fork, hipEvent_t join, KernelProbe *probe = nullptr);

// how much deeper than the last pass's frame perimeter pass j (of nmid + 1)
// reaches into the interior: 2 nodes for the pass before the last (the next
// pass's halo computes a first step one node in, from pulls two nodes in), 2
// deeper for each pass before that
__host__ __device__ constexpr int ring_depth(int nmid, int j) { return 2 + 2 * (nmid - 1 - j); }

// The frame of a deep march (LEVELS = 2 (nmid + 1) steps, H = LEVELS
// - 1): its interior waves cover rows [H + 2, ly - H) and columns
// [H + 2, lx - H); perimeter pass j (steps 2j + 1, 2j + 2) covers the
// frame around them widened inwards by the depth t_j its successor reads
// (t_nmid = 0, t_j = t_{j+1} + 2: the next pass's halo computes a first step
// one node further in, from pulls two nodes in), so no other kernel is needed:
// strips of rows [1, 1 + sr_j) and [ly + 1 - sr_j, ly + 1) over the whole
// width (chunks of the strip grid sg, 64 / gl per wave, gl = 8 lanes per
// chunk for sr_j + 2 <= 8, else 16), and the columns
// [1, 1 + sr_j) and [lx + 1 - sr_j, lx + 1) of the rows between them in
// 62-row bands, with sr_j = H + 1 + t_j.
struct DeepFrame {
    int nmid;            // passes - 1
    int sr[4];           // strip rows (= side columns) of each pass
    int nsw[4];          // waves per strip of each pass
    int nsb[4];          // side bands of each pass, per side
    WaveGrid sg;         // the strips' chunks (cx, ncx; right-aligned at lx + 1)
};

// Many channels of a batch (pool offset `off` elements), several steps per
// launch with the register marches over each channel's wave grid.  The
// channels' P and Q sit at A + off and B + off.  A work table lists
// (channel, wave) for perimeter waves and (channel, first wave / WPG) for the
// workgroups of interior waves (the workgroups of a channel contiguous).
struct ChannelWaves {
    Problem p;
    WaveGrid g;
    DeepFrame f;     // deep grids only
    int64_t off;
};
struct BatchTables {
    const ChannelWaves *ch;
    const int2 *edge;
    int nedge;
    const int2 *in;
    int nin_wgs;
};
// Two steps: perimeter waves on `side`, interior waves on `st`, joined back
// into `st`; dstep: index of the first collision relative to the counter.
int batch_wpanel_wpg(int storage);
hipError_t launch_batch_wpanel(int storage, const BatchTables &t, uint32_t dstep, const void *A,
                               void *B, hipStream_t st, hipStream_t side, hipEvent_t fork,
                               hipEvent_t join);
// LEVELS = 4 or 6 steps: the channels with a deep grid (deep: its interior
// waves on `st`; its perimeter passes, frame_items[j] (channel, wave), on
// `side` through mid[0 .. LEVELS/2 - 2]) and the others (shallow, on their
// two-step grids: LEVELS/2 two-step launches on `side` through the same
// buffers).  batch_deep_grid: false when the channel is too small for it.
bool batch_deep_grid(const Layout &L, int cx, int levels, WaveGrid *g, DeepFrame *f);
// the grid and frame of a deep march of LEVELS = 4, 6 or 8 steps in cx-column
// chunks, strips in strip_cx-column chunks (false: the lattice is too small)
bool aligned_deep_grid(const Layout &L, int cx, int strip_cx, int levels, WaveGrid *g,
                       DeepFrame *f);
int batch_frame_waves(const DeepFrame &f, int j);
// the strip chunk width of a single lattice's deep march in cx-column chunks
int deep_strip_cx(int cx);
int batch_deep_wpg(int storage, int levels);
hipError_t launch_batch_wpanel_deep(int storage, int levels, const BatchTables &deep,
                                    const int2 *const *frame_items, const int *nframe,
                                    const BatchTables &shallow, const void *A, void *const *mid,
                                    void *B, hipStream_t st, hipStream_t side, hipEvent_t fork,
                                    hipEvent_t join);

}  // namespace lbm
