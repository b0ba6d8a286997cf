// kernels_wave.hip -- several time steps per launch with the intermediate
// steps held in registers of one wave64: the panel sweep
// (collide_tight_panel_iy, lb.c:1402-1538, two steps) mapped onto CDNA
// wavefronts, and the same march four or six steps deep (no counterpart in
// the reference).
//
// A wave owns a band of rows and a chunk of columns, and marches x.  For two
// steps its 64 lanes are the band's 62 rows plus one halo row on each side
// (lane l <-> row Y0-1+l).  Iteration x computes the first step (m) of
// column x for all 64 rows into registers, then the second step (m+1) of
// column x-1 for the 62 owned rows.  The second step needs the first step's
// populations of columns x-2, x-1 and x at rows y-1, y, y+1: a lane keeps its
// own rows' values of the last two columns in registers (column x-1: all 9
// populations; x-2: populations 1, 5, 8, the cx = +1 sources) and takes its
// neighbours' with one lane shift up and one down (6 values per column).
// Nothing goes through LDS.  Four steps stack three such levels (58 owned
// rows, three halo rows each side), six steps five (54 owned rows).  The
// deep sweeps' perimeter frame advances two steps per pass through further
// buffers, each pass but the last widened inwards by what the next one reads.
//
// Waves whose rows or columns touch the channel perimeter run a separate,
// general kernel (k_wpanel_edge: wall operators, outlet extrapolation,
// unknown populations; the strips of a few rows along the walls take
// several chunks per wave); the interior kernel (k_wpanel_in) runs BGK /
// bounce-back only, branch-free.  The second step's outlet density is
// extrapolated from the post-stream densities of columns lx-2 and lx-1
// gathered by the same lane just before their second collision
// (lb.c:968-977); chunks are right-aligned so the chunk holding lx also holds
// lx-2 and lx-1.  Unknown populations and the ghost ring follow the
// collided-state convention of kernels_f64.hip: the first step pulls them
// from P's ghost ring; the second step's are P(x) itself; the output Q's
// ghost ring gets the first step's values.
//
// Storage S = double (bit-exact FP64) or float / _Float16 deviations (FP32:
// FP64 collisions rounded stochastically; FP16: FP32 collisions; FP64 walls).
// For reduced storage each intermediate step is rounded to S exactly as a
// store would round it, so k steps here equal k single steps bit for bit.
#include <cstdio>
#include <utility>

#include "lanes.hpp"
#include "launch.hpp"
#include "step_any.hpp"

namespace lbm {

// ---- perimeter waves ---------------------------------------------------------
// Waves whose rows or columns (halo included) touch the channel perimeter
// carry the wall operators, the outlet extrapolation and the unknown-population
// bookkeeping.  One wave per workgroup; task e enumerates the frame of
// (band, chunk) pairs around the interior rectangle.
__device__ __forceinline__ void edge_task(const WaveGrid &g, int e, int &band, int &ck) {
    const int nfull = g.b_lo + (g.nbands - g.b_hi);     // bands with every chunk on the edge
    if (e < nfull * g.ncx) {
        int i = e / g.ncx;
        band = i < g.b_lo ? i : g.b_hi + (i - g.b_lo);
        ck = e % g.ncx;
    } else {
        e -= nfull * g.ncx;
        int ns = g.c_lo + (g.ncx - g.c_hi);              // edge chunks of an interior band
        band = g.b_lo + e / ns;
        int j = e % ns;
        ck = j < g.c_lo ? j : g.c_hi + (j - g.c_lo);
    }
}

// One perimeter wave's two-step march over its chunk(s).  GL < 64: groups of
// GL lanes, group k on chunk ck + k of a strip of at most GL - 2 rows (rows
// Y0 .. Y0+nrows-1 on lanes 1 .. nrows of the group, halo rows on lanes 0 and
// nrows+1: a group's owned rows read shifted values from their own group
// only); GL = 64: the columns [xa, xb), rows Y0 .. Y0+nrows-1 on lanes
// 1 .. nrows.  dstep: index of this launch's first collision relative to
// the counter.
template <typename S, int GL>
__device__ __forceinline__ void edge_march(const Problem &p, const WaveGrid &g, uint32_t dstep,
                                           const S *__restrict__ P, S *__restrict__ Q, int Y0,
                                           int nrows, int ck, int xa = 0, int xb = 0) {
    using T = typename Calc<S>::T;
    constexpr bool THIN = GL < 64;
    const Layout &L = p.lay;
    if constexpr (sizeof(S) == 4) {
        pin_index(Y0);
        pin_index(nrows);
        pin_index(ck);
    }
    const int lane = threadIdx.x;
    const int lr = lane & (GL - 1);                      // row Y0 - 1 + lr
    const int ckl = ck + lane / GL;
    const bool active = !THIN || ckl < g.ncx;            // THIN: a strip's last wave may run short
    const WaveSpan w = wave_span(L, g, 0, active ? ckl : g.ncx - 1);
    const int X0 = THIN ? w.X0 : xa, X1 = THIN ? w.X1 : xb;
    const int y = Y0 - 1 + lr;
    const bool row_in = active && y >= 1 && y <= L.ly;
    const bool owner = active && lr >= 1 && lr <= nrows && y <= L.ly;
    const int s_row = (y == 1) ? LOWER : (y == L.ly) ? UPPER : -1;
    const uint32_t m = sizeof(S) == 4 ? *p.step + dstep : 0u;   // index of the first collision

    T cur[9], c1[9], c2[3];      // columns x, x-1 and (1, 5, 8 of) x-2
    S fn[9];                     // next column's pulled populations
    for (int i = 0; i < 9; ++i) { cur[i] = 0; c1[i] = 0; fn[i] = 0; }
    c2[0] = c2[1] = c2[2] = 0;
    double rho1 = 0., rho2 = 0.;

    auto pull = [&](int x, S f[9]) {
        int64_t n = L.node(x, y);
        for_pops([&]<int I>() { f[I] = P[I * L.plane + n - cx(I) * L.pitch - cy(I)]; });
    };
    if (row_in && X0 - 1 >= 1) pull(X0 - 1, fn);
    if constexpr (sizeof(S) == 4) fp32_round_toward_zero();

    auto column = [&](int x) {
        // ---- step m, column x, all 64 rows -> registers
        for_pops([&]<int I>() { cur[I] = as_calc<S>(fn[I]); });
        if (row_in && x + 1 >= 1 && x + 1 <= X1 && x + 1 <= L.lx) pull(x + 1, fn);
        if (row_in && x >= 1 && x <= L.lx) {
            int s = s_row >= 0 ? s_row : (x == 1 ? LEFT : (x == L.lx ? RIGHT : -1));
            double rp0 = 0.;
            if (s == RIGHT && p.zgb)
                rp0 = zgb_rho(pulled_rho<S>(L, P, L.lx - 1, y), pulled_rho<S>(L, P, L.lx - 2, y));
            collide_any<S>(p, s, x, y, cur, rp0, m);
            for_pops([&]<int I>() { cur[I] = stored<S>(cur[I]); });
        }
        // ---- neighbour rows' first-step values (all lanes take part)
        const T up2 = lane_up(c1[2]), up5 = lane_up(c2[1]), up6 = lane_up(cur[6]);
        const T dn4 = lane_down(c1[4]), dn7 = lane_down(cur[7]), dn8 = lane_down(c2[2]);
        // ---- step m+1, column x-1, owned rows
        const int xo = x - 1;
        if (owner && xo >= X0) {
            T h[9] = {c1[0], c2[0], up2, cur[3], dn4, up5, up6, dn7, dn8};
            int64_t n = L.node(xo, y);
            int s = s_row >= 0 ? s_row : (xo == 1 ? LEFT : (xo == L.lx ? RIGHT : -1));
            if (s >= 0) {
                // populations entering from outside: this node's values of
                // step m-1 (P itself); Q's ghost ring gets step m's
                for_pops([&]<int I>() {
                    if constexpr (I != 0) {
                        int gx = xo - cx(I), gy = y - cy(I);
                        if (outside(L, gx, gy)) {
                            h[I] = as_calc<S>(P[I * L.plane + n]);
                            Q[I * L.plane + L.node(gx, gy)] = (S)c1[I];
                        }
                    }
                });
            }
            if (xo == L.lx - 2) rho2 = rho_of_any<S>(h);
            if (xo == L.lx - 1) rho1 = rho_of_any<S>(h);
            double rp0 = 0.;
            if (s == RIGHT) {
                rp0 = zgb_rho(rho1, rho2);
                p.outlet_rho[y] = p.zgb ? rp0 : p.side[RIGHT].p0[y];
            }
            collide_any<S>(p, s, xo, y, h, rp0, m + 1);
            for_pops([&]<int I>() {
                if constexpr (sizeof(S) == 8) __builtin_nontemporal_store(h[I], &Q[I * L.plane + n]);
                else __builtin_nontemporal_store(narrow<S>(h[I]), &Q[I * L.plane + n]);
            });
        }
        c2[0] = c1[1];
        c2[1] = c1[5];
        c2[2] = c1[8];
        for_pops([&]<int I>() { c1[I] = cur[I]; });
    };
    if constexpr (THIN) {
        // the same cx + 2 iterations for every group: a group's march starts
        // at X1 - cx - 1, below X0 - 1 for a partial first chunk, whose extra
        // iterations neither load nor compute
        const int xbeg = X1 - g.cx - 1;
        for (int k = 0; k <= g.cx + 1; ++k) column(xbeg + k);
    } else {
        for (int x = X0 - 1; x <= X1; ++x) column(x);
    }
}

// Perimeter wave e of grid g.  GL: lanes per chunk of the thin strips
// (g.gl; 4 for the two-step grid's two-row strips, 8 for the four- and
// six-step grids' strips of up to six rows, 16 for the eight-step grid's
// eight)
template <typename S, int GL>
__device__ __forceinline__ void wpanel_edge_wave(const Problem &p, const WaveGrid &g, int e,
                                                 uint32_t dstep, const S *__restrict__ P,
                                                 S *__restrict__ Q) {
    if (g.thin) {
        // waves [0, nthin): the strips below r0 and from r1 up, 64 / GL
        // chunks each; then the edge chunks of rows [r0, r1) in 62-row bands
        const int nps = g.nthin / 2;
        if (e < g.nthin) {
            const bool top = e >= nps;
            const int ck0 = (top ? e - nps : e) * (64 / GL);
            edge_march<S, GL>(p, g, dstep, P, Q, top ? g.r1 : 1,
                              top ? p.lay.ly + 1 - g.r1 : g.r0 - 1, ck0);
            return;
        }
        e -= g.nthin;
        const int ns = g.c_lo + (g.ncx - g.c_hi);
        const int i = e / ns, j = e % ns;
        const int Y0 = g.r0 + i * WAVE_ROWS;
        const WaveSpan w = wave_span(p.lay, g, 0, j < g.c_lo ? j : g.c_hi + (j - g.c_lo));
        edge_march<S, 64>(p, g, dstep, P, Q, Y0, g.r1 - Y0 < WAVE_ROWS ? g.r1 - Y0 : WAVE_ROWS,
                          0, w.X0, w.X1);
        return;
    }
    int band, ck;
    edge_task(g, e, band, ck);
    const WaveSpan w = wave_span(p.lay, g, 0, ck);
    edge_march<S, 64>(p, g, dstep, P, Q, 1 + band * WAVE_ROWS, WAVE_ROWS, 0, w.X0, w.X1);
}

template <typename S, int GL>
__global__ __launch_bounds__(64) void k_wpanel_edge(Problem p, WaveGrid g, uint32_t dstep,
                                                    const S *__restrict__ P, S *__restrict__ Q) {
    wpanel_edge_wave<S, GL>(p, g, (int)blockIdx.x, dstep, P, Q);
}

// ---- interior waves -------------------------------------------------------------
// Every lane is a channel row away from the walls and every column of the
// march is an interior column, so the march is written without branches
// around memory operations: each iteration issues the same loads and stores.
// Lanes that must not store (the two halo rows, and every lane during the
// two pipeline-filling iterations) get a buffer offset past the end of the
// buffer resource, and the hardware drops their stores.  With identical
// memory traffic on every path the compiler's wait for a column's loads
// counts only the operations issued after them, so the stores are never
// waited for and the loads of DEPTH columns stay in flight behind the
// arithmetic.
constexpr uint32_t DROP = 0x80000000u;        // offset that no resource here reaches
constexpr int RSRC_WORD3 = 0x00020000;        // raw buffer descriptor word 3 (CDNA)
// Stores go through L2 with the default (write-back) policy, not as
// non-temporal streaming stores: a wave's owned rows (62 or 58 of them) do
// not fill whole 128-byte lines, and a line shared by two waves is only
// complete once both have written it.  Kept in L2, the two halves merge
// before the line goes to HBM; streamed, each half reaches HBM as a partial
// write.  On 16384 x 4096 the two-step FP64 sweep runs 5% faster this way,
// FP32 13%, FP16 20%.
constexpr int CPOL_STORE = 0;

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

template <typename S>
__device__ __forceinline__ void buffer_store(S v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(S) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, CPOL_STORE);
    else if constexpr (sizeof(S) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, 0, CPOL_STORE);
    else
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), r, off, 0, CPOL_STORE);
}

// Compile-time loop over K = 0 .. N-1.
template <int N, typename F> __device__ __forceinline__ void static_for(F &&f) {
    [&]<int... K>(std::integer_sequence<int, K...>) __attribute__((always_inline)) {
        (f.template operator()<K>(), ...);
    }(std::make_integer_sequence<int, N>{});
}

// WPG waves per workgroup take consecutive bands and march in lockstep (one
// barrier per column), so the workgroup's loads of a column arrive at the
// memory together as one contiguous run per population: eight waves.  The
// six-step FP64 march (205 VGPRs with two columns in flight, two waves per
// SIMD) takes four: an eight-wave workgroup needs two waves on every SIMD of
// its CU, so no perimeter wave (242 VGPRs in that measurement; the frame
// kernel k_frame takes 184) fits beside it, while beside a four-wave
// workgroup one does (16384 x 4096: 129.1-129.6 GLUPS against
// 116.2-116.9, and 115.0 / 119.2 with two / one; the four-step sweep, more
// memory-bound, loses with four: 95 against 113).  The six-step march with
// one column in flight takes four as well (twelve would hold it to 168
// VGPRs, where it spills); so does the eight-step FP64 march (252 VGPRs, one
// column in flight: 3.69-3.74 ms per launch with four, 4.14 with eight).  A
// six-step variant at three waves per SIMD (the
// next level's c_x = +1 window in LDS, 164 VGPRs, one column in flight) ran
// its interior kernel in 2.53 ms against 2.72, but then no perimeter wave
// fits beside it until it drains, and the launch took 3.25 ms (2.70 with the
// perimeter kernel held to 168 VGPRs and spilling).  With one column in
// flight (126 VGPRs, three waves per SIMD) the FP64 four-step march takes
// twelve, the whole CU (115.0 against 112.6 at 63 columns); so do the FP64
// two-step march (63.5 against 61.6) and the FP16 four-step one (239.0
// against 228.1).  The FP32 six-step march also takes four (134.7 against
// 123.3; FP16 loses with four: 176.8 against 225.8).  Measured and left at
// eight: FP32 (panel 119.9 / 120.1, panel4 129.2 / 138.5 with twelve /
// eight), FP16 panel (212.7 / 211.7).
template <typename S, int DEPTH, int LEVELS> constexpr int wpg_of() {
    if constexpr (sizeof(S) >= 4 && LEVELS >= 6) return 4;
    else if constexpr (sizeof(S) == 8 && LEVELS >= 4 && DEPTH == 1) return 12;
    else if constexpr (sizeof(S) == 8 && LEVELS == 2 && DEPTH == 2) return 12;
    else if constexpr (sizeof(S) == 2 && LEVELS == 4 && DEPTH == 2) return 12;
    else return 8;
}

// Interior waves, LEVELS time steps per launch (2, or 4 / 6 for the deeper
// sweeps).  A wave's 64 lanes are rows Y0-H .. Y0+63-H with H = LEVELS-1 halo
// rows on each side: level l (1-based) is valid on lanes l-1 .. 64-l, so the
// wave owns the 64-2H rows of the last level.  Iteration x computes level 1
// of column x from the pulled populations and level l of column x-l+1 from
// level l-1's window of three columns (two kept in registers per level, the
// third just computed), then stores level LEVELS of column x-H.
// DEPTH = columns whose loads are in flight (1 or 2); cx % DEPTH == 0.
// dstep: index of the launch's first collision relative to the counter.
// Interior wave t of grid g (t >= g.nin: nothing, the whole wave leaves);
// the WPG waves of a workgroup march in lockstep, so they take interior
// waves of one grid.
// DISC: the wave's nodes may reach the disc (per-lane bounce-back branch).
// Without that branch an iteration is one basic block, and the scheduler
// interleaves its collisions so freely that the six-step FP64 march needs
// 256 registers and spills (1.3x slower); scheduling fences between the
// collisions (fence_after) keep the branchy form's register use (205 VGPRs,
// two waves per SIMD) and still drop its branches and the copies at their
// merges.
// Fences that let memory or scalar instructions cross measured the same;
// unrolling the march by six columns (no window copies at all, 13% fewer
// instructions per column) 1.5% faster for FP64 and 15% slower for FP32.

// scheduling fence after the collisions of levels 1, 4, 7 (every collision
// fenced: two-step march 64.5 GLUPS, four-step 117.5; fenced after levels 1,
// 4, 7: 68.3 and 119.3; six-step unchanged at 205 VGPRs either way)
constexpr bool fence_after(int level) { return level % 3 == 1; }

template <typename S, int DEPTH, int LEVELS, bool DISC>
__device__ __forceinline__ void wpanel_in_march(const Problem &p, const WaveGrid &g, const int t,
                                                uint32_t dstep, const S *__restrict__ P,
                                                S *__restrict__ Q) {
    using T = typename Calc<S>::T;
    constexpr int H = LEVELS - 1;
    constexpr int ROWS = 64 - 2 * H;
    const Layout &L = p.lay;
    const int lane = threadIdx.x & 63;
    if (t >= g.nin) return;                              // whole waves only
    const int ck = g.c_lo + t / g.nbl;
    const int Y0 = g.r0 + (t % g.nbl) * ROWS;
    const int X1 = in_chunk_x1(g, ck);                  // interior chunks are full
    const int X0 = X1 - g.cx;
    const int y = Y0 - H + lane;
    // a partial last band: lanes past the interior's halo read row ly (in
    // bounds; their values reach no owned row)
    const int yl = y < L.ly ? y : L.ly;
    const uint32_t m = sizeof(S) == 4 ? *p.step + dstep : 0u;   // first collision's index
    const uint32_t colb = (uint32_t)L.pitch * (uint32_t)sizeof(S);
    const uint32_t lane_off = (lane >= H && lane < H + ROWS && y < g.r1)
                                  ? (uint32_t)(y + L.yoff) * (uint32_t)sizeof(S) : DROP;
    // per population: columns [X0, X1) of Q
    __amdgpu_buffer_rsrc_t rs[9];
    for_pops([&]<int I>() {
        rs[I] = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(Q + I * L.plane + (int64_t)X0 * L.pitch), 0,
            (int)((uint32_t)(X1 - X0) * colb), RSRC_WORD3);
    });

    S f[DEPTH][9];
    // per level 1..LEVELS-1: columns x-l+1 (cur), x-l (c1) and populations
    // 1, 5, 8 of x-l-1 (c2); index 0 unused
    T cur[LEVELS][9], c1[LEVELS][9], c2[LEVELS][3];
    for (int d = 0; d < DEPTH; ++d)
        for (int i = 0; i < 9; ++i) f[d][i] = 0;
    for (int l = 0; l < LEVELS; ++l) {
        for (int i = 0; i < 9; ++i) { cur[l][i] = 0; c1[l][i] = 0; }
        c2[l][0] = c2[l][1] = c2[l][2] = 0;
    }

    // per population, a wave-uniform base (SGPRs) with pb[I][v(x)] the value
    // pulled into (x, y), v(x) = (x - X0 + H) * pitch + y + yoff: one 32-bit
    // lane offset per column serves all nine loads
    const S *pb[9];
    for_pops([&]<int I>() {
        pb[I] = P + I * L.plane + (int64_t)(X0 - H - cx(I)) * L.pitch - cy(I);
    });
    const uint32_t vy = (uint32_t)(yl + L.yoff) * (uint32_t)sizeof(S);
    auto pull = [&](int x, S (&dst)[9]) {
        // byte offset kept 32-bit so the loads take the SGPR-base form
        const uint32_t v = (uint32_t)(x - X0 + H) * colb + vy;
        for_pops([&]<int I>() {
            dst[I] = *reinterpret_cast<const S *>(reinterpret_cast<const char *>(pb[I]) + v);
        });
    };
    // One iteration: level 1 of column x, then levels 2 .. TOP (TOP < LEVELS
    // in the head iterations, where the higher levels' columns lie outside
    // what the chunk needs), the store of level LEVELS (dropped unless
    // TOP == LEVELS), and the window shifts.  FILL iterations only put loads
    // in flight; TAIL iterations are the last DEPTH, with nothing left to
    // load.  Every kind issues the same stores, so the memory traffic of the
    // loop and of the code around it looks alike to the compiler's wait-count
    // analysis.
    enum { FILL, HEAD, FULL, TAIL };
    auto iter = [&]<int KIND, int TOP>(int x, S (&buf)[9]) __attribute__((always_inline)) {
        if constexpr (KIND == FILL) {
            pull(x + DEPTH, buf);
            for_pops([&]<int I>() { buffer_store<S>(S(0), rs[I], DROP); });
        } else {
            for_pops([&]<int I>() { cur[1][I] = as_calc<S>(buf[I]); });
            if constexpr (KIND != TAIL) pull(x + DEPTH, buf);
            collide_bulk<S, DISC>(p, x, y, cur[1], m);
            if constexpr (!DISC && fence_after(1)) __builtin_amdgcn_sched_barrier(0);
            for_pops([&]<int I>() { cur[1][I] = stored<S>(cur[1][I]); });
            static_for<TOP - 1>([&]<int K>() __attribute__((always_inline)) {
                constexpr int l = K + 2;                 // level computed, column x-l+1
                constexpr int b = l - 1;                 // the level it gathers from
                const T up2 = lane_up(c1[b][2]), up5 = lane_up(c2[b][1]), up6 = lane_up(cur[b][6]);
                const T dn4 = lane_down(c1[b][4]), dn7 = lane_down(cur[b][7]), dn8 = lane_down(c2[b][2]);
                T h[9] = {c1[b][0], c2[b][0], up2, cur[b][3], dn4, up5, up6, dn7, dn8};
                collide_bulk<S, DISC>(p, x - l + 1, y, h, m + l - 1);
                if constexpr (!DISC && fence_after(l)) __builtin_amdgcn_sched_barrier(0);
                if constexpr (l < LEVELS) {
                    for_pops([&]<int I>() { cur[l][I] = stored<S>(h[I]); });
                } else {
                    const uint32_t off = lane_off + (uint32_t)(x - H - X0) * colb;
                    for_pops([&]<int I>() { buffer_store<S>(to_storage<S>(h[I]), rs[I], off); });
                }
            });
            if constexpr (TOP < LEVELS)
                for_pops([&]<int I>() { buffer_store<S>(S(0), rs[I], DROP); });
            // shift the windows of the levels computed in this iteration
            static_for<(TOP < LEVELS ? TOP : LEVELS - 1)>([&]<int K>() __attribute__((always_inline)) {
                constexpr int l = K + 1;
                c2[l][0] = c1[l][1];
                c2[l][1] = c1[l][5];
                c2[l][2] = c1[l][8];
                for_pops([&]<int I>() { c1[l][I] = cur[l][I]; });
            });
        }
        __builtin_amdgcn_s_barrier();
    };

    // x runs X0-H-DEPTH .. X1+H-1: DEPTH fill iterations, 2H head iterations
    // (head i computes levels up to i/2 + 1), FULL up to X1+H-1-DEPTH, then
    // DEPTH tail iterations; the j-th iteration (from 0) consumes and reloads
    // buffer j % DEPTH
    int x = X0 - H - DEPTH;
    if constexpr (sizeof(S) == 4) fp32_round_toward_zero();
    static_for<DEPTH>([&]<int D>() __attribute__((always_inline)) {
        iter.template operator()<FILL, 1>(x + D, f[D]);
    });
    x += DEPTH;
    static_for<2 * H>([&]<int I>() __attribute__((always_inline)) {
        iter.template operator()<HEAD, I / 2 + 1>(x + I, f[I % DEPTH]);
    });
    x += 2 * H;
    for (int k = 0; k < (X1 - X0) / DEPTH - 1; ++k, x += DEPTH)
        static_for<DEPTH>([&]<int D>() __attribute__((always_inline)) {
            iter.template operator()<FULL, LEVELS>(x + D, f[(2 * H + D) % DEPTH]);
        });
    static_for<DEPTH>([&]<int D>() __attribute__((always_inline)) {
        iter.template operator()<TAIL, LEVELS>(x + D, f[(2 * H + D) % DEPTH]);
    });
}

// The march of interior wave t, with the bounce-back test only in waves
// whose nodes (halo rows and columns included) can reach the disc: the
// others, nearly all, collide without a per-lane branch (one basic block per
// column, no merge copies; on 16384 x 4096 with the shared-denominator
// division of macros() the six-step FP64 march runs 3.00 -> 2.73 ms per
// launch, the four-step 2.31 -> 2.25, FP32 six-step 3.0 -> 2.9 ms).  The
// test is wave-uniform, and both forms issue the same barriers, so a
// workgroup's waves stay in lockstep.
template <typename S, int DEPTH, int LEVELS>
__device__ __forceinline__ void wpanel_in_wave(const Problem &p, const WaveGrid &g, const int t,
                                               uint32_t dstep, const S *__restrict__ P,
                                               S *__restrict__ Q) {
    constexpr int H = LEVELS - 1;
    if (t >= g.nin) return;
    const int ck = g.c_lo + t / g.nbl;
    const int Y0 = g.r0 + (t % g.nbl) * (64 - 2 * H);
    const int X1 = in_chunk_x1(g, ck);
    const int X0 = X1 - g.cx;
    // FP16 keeps the branchy form everywhere: its split marches measured
    // slower (two-step -2%, four-step -16%, six-step -9%)
    if (sizeof(S) == 2 || box_meets_disc(p, X0 - H - DEPTH, X1 + H, Y0 - H, Y0 - H + 63))
        wpanel_in_march<S, DEPTH, LEVELS, true>(p, g, t, dstep, P, Q);
    else
        wpanel_in_march<S, DEPTH, LEVELS, false>(p, g, t, dstep, P, Q);
}

template <typename S, int DEPTH, int LEVELS>
__global__ __launch_bounds__((64 * wpg_of<S, DEPTH, LEVELS>())) void k_wpanel_in(
    Problem p, WaveGrid g, uint32_t dstep, const S *__restrict__ P, S *__restrict__ Q) {
    constexpr int WPG = wpg_of<S, DEPTH, LEVELS>();
    const int t = __builtin_amdgcn_readfirstlane(
        xcd_block(blockIdx.x, gridDim.x) * WPG + (threadIdx.x >> 6));
    wpanel_in_wave<S, DEPTH, LEVELS>(p, g, t, dstep, P, Q);
}

// Chunk width, unless given: the widest even width in [4, 16] that still
// yields about 16384 waves, so mid-size lattices get enough waves in flight
// (2048 x 512 FP64: 38.7 GLUPS at 4 columns, 22.3 at 16) while large ones
// keep the (cx + 2) / cx first-step overhead low (16384 x 4096: 59 GLUPS at
// 16, 58 at 32, 55 at 128).
static WaveGrid wave_grid_max(const Layout &L, int cx, int cmax, int64_t waves) {
    WaveGrid g;
    if (cx < 3) {
        const int64_t bands = (L.ly + WAVE_ROWS - 1) / WAVE_ROWS;
        const int64_t c = bands * L.lx / waves;
        cx = (int)(c < 4 ? 4 : c > cmax ? cmax : c) & ~1;
    }
    g.cx = cx;
    // the interior kernel's buffer resources span one chunk's columns and its
    // 32-bit load offsets the chunk plus its halo columns (at most 2 x 7 + 2
    // more): both stay below the drop offset (2 GiB) at up to 8 bytes per
    // element (lattices too tall for 3 columns: wave_grid_fits)
    while (g.cx > 3 && (int64_t)(g.cx + 16) * L.pitch * 8 >= (int64_t)DROP) g.cx /= 2;
    g.nbands = (L.ly + WAVE_ROWS - 1) / WAVE_ROWS;
    g.ncx = (L.lx + g.cx - 1) / g.cx;
    g.ntasks = g.nbands * g.ncx;
    // a wave is interior when its rows (halo included) stay off rows 1 and ly
    // and its columns off columns 1 and lx: a band condition and a chunk one
    int b_lo = g.nbands, b_hi = 0, c_lo = g.ncx, c_hi = 0;
    for (int b = 0; b < g.nbands; ++b) {
        const WaveSpan w = wave_span(L, g, b, 0);
        if (w.Y0 - 1 > 1 && w.Y0 + WAVE_ROWS < L.ly) { b_lo = b < b_lo ? b : b_lo; b_hi = b + 1; }
    }
    for (int c = 0; c < g.ncx; ++c) {
        const WaveSpan w = wave_span(L, g, 0, c);
        if (w.X0 - 1 > 1 && w.X1 < L.lx) { c_lo = c < c_lo ? c : c_lo; c_hi = c + 1; }
    }
    if (b_lo >= b_hi || c_lo >= c_hi) b_lo = b_hi = c_lo = c_hi = 0;
    g.b_lo = b_lo; g.b_hi = b_hi; g.c_lo = c_lo; g.c_hi = c_hi;
    g.nin = (b_hi - b_lo) * (c_hi - c_lo);
    g.nedge = g.ntasks - g.nin;
    g.r0 = 1 + b_lo * WAVE_ROWS;
    g.r1 = 1 + b_hi * WAVE_ROWS;
    g.nbl = b_hi - b_lo;
    g.thin = g.nthin = g.gl = 0;
    g.xr = L.lx + 1;
    g.xl = 1;
    return g;
}

// The thin frame of a grid whose interior waves keep an H-row / H-column
// halo: interior chunks shrunk until the halo stays off columns 1 and lx,
// interior rows r0 = H + 2 .. r1 = ly - H (the rows whose halo stays within
// rows 2 .. ly - 1) in (64 - 2H)-row bands, the last one partial.  The frame
// keeps H + 1 rows along each wall, taken 64 / gl chunks per perimeter wave
// (gl = 4 lanes per chunk for H = 1, 8 for H = 3 and 5), and the edge chunks
// of rows [r0, r1) in 62-row bands.  Fewer than
// min_rows interior rows or no interior chunk: every wave is a perimeter
// wave.
static WaveGrid make_thin(const Layout &L, WaveGrid g, int H, int min_rows) {
    const int rows = 64 - 2 * H;
    auto chunk_x0 = [&](int ck) { return L.lx + 1 - (g.ncx - 1 - ck) * g.cx - g.cx; };
    // interior chunks by the halo condition alone (the band test of the
    // two-step grid finds none on lattices below 126 rows)
    g.c_lo = 0;
    g.c_hi = g.ncx;
    while (g.c_lo < g.c_hi && chunk_x0(g.c_lo) - H < 2) ++g.c_lo;
    while (g.c_hi > g.c_lo && chunk_x0(g.c_hi - 1) + g.cx - 1 + H > L.lx - 1) --g.c_hi;
    g.r0 = 2 + H;
    g.r1 = L.ly - H;
    if (g.c_lo >= g.c_hi || g.r1 - g.r0 < min_rows) {
        g.b_lo = g.b_hi = g.c_lo = g.c_hi = 0;
        g.nin = 0;
        g.nedge = g.ntasks;
        g.nbl = 0;
        g.thin = g.nthin = g.gl = 0;
        return g;
    }
    g.thin = 1;
    g.b_lo = g.b_hi = 0;   // the thin frame has no (band, chunk) rectangle
    g.nbl = (g.r1 - g.r0 + rows - 1) / rows;
    g.nin = g.nbl * (g.c_hi - g.c_lo);
    g.gl = H + 3 <= 4 ? 4 : H + 3 <= 8 ? 8 : 16;   // H + 1 strip rows and two halo rows per chunk
    const int cpw = 64 / g.gl;
    g.nthin = 2 * ((g.ncx + cpw - 1) / cpw);
    g.nedge = g.nthin + (g.r1 - g.r0 + WAVE_ROWS - 1) / WAVE_ROWS * (g.c_lo + g.ncx - g.c_hi);
    return g;
}

// The two-step grid: an interior of at least one band of rows (below that
// the interior kernel's launch costs more than its rows save: 256 x 64 ran
// 1.2 GLUPS with a 60-row interior against 1.8 on perimeter waves alone).
WaveGrid wave_grid(const Layout &L, int cx) {
    return make_thin(L, wave_grid_max(L, cx, 16, 16384), 1, WAVE_ROWS);
}

// The thin-frame grid of the deep sweeps (LEVELS = 4, 6 or 8 steps per
// launch, H = LEVELS-1 halo rows and columns): the interior chunks of the
// two-step grid, shrunk until their H-column halo stays off columns 1 and lx,
// over rows r0 = H + 2 .. r1 = ly - H, tiled by (64 - 2H)-row bands.  It
// decides whether a lattice takes a deep sweep (wpanel_deep_possible) and the
// chunk width; the launch then runs on the aligned grid over the same rows
// (aligned_deep_grid).  The deep
// march recomputes 2H halo columns per chunk and does LEVELS/2 times the
// two-step arithmetic per column, so it wants wider chunks, up to 24, and
// fewer waves (about 6144) before narrowing them: four-step FP64 GLUPS at 4 /
// 8 / 12 / 24 columns: 4096 x 1024 56 / 65 / 66 / 51, 8192 x 2048 66 / 80 /
// 85 / 86, 16384 x 4096 - / 91 / 97 / 100.  Only the largest lattices gain
// from chunks wider still (fewer halo columns re-read): 16384 x 4096 at 24 /
// 40 / 64 columns runs FP64 104 / 106 / 107, FP32 128 / 134 / 132, but FP16
// 217 / 204 / 187 (frame strips as wide as the interior chunks, as below),
// and 8192 x 2048 loses at 40 in every format (FP32 105 at
// 24, 91 at 40); four steps take 40 columns in FP64 and FP32 where 24 would
// leave more than 128 x 6144 waves.  There, FP64 runs one column of loads in
// flight better than two (odd widths: 39 / 47 / 55 / 63 columns 106.7 /
// 107.2 / 107.5 / 107.7 GLUPS against 105.7 / 105.6 / 104.9 / 106.0 at 40 /
// 48 / 56 / 64) and takes 63.  Six steps (ten halo columns per chunk) take
// 64 (FP64 101 / 101 / 106 at 47 / 55 / 64), and FP64 96 once the frame's
// strips run in FRAME_CX-column chunks: the six-step launch at 64 / 96 / 128
// / 192 columns 149.4-150.0 / 150.7-151.8 / 150.4-152.1 / 146.2-147.8 GLUPS,
// the 6 + 6 + 8 plan 147.1-149.5 / 150.0-151.3 / 148.9-149.6 / 147.6-148.3;
// the eight-step FP64 launch too: 3.73-3.79 ms at 64 columns, 3.60-3.78 at
// 96, 3.73-3.80 at 128, 3.77-3.80 at 160 (six-step launches at 96).  With
// FRAME_CX-column strips the reduced formats' four-step sweeps take 96 as
// well: FP16 231-235 GLUPS at 24 columns, 238-241 at 40, 246-250 at 96;
// FP32 140 at 40, 143-144 at 96 (FP64 at 95: 123.3-123.5, as at 63); so do
// their six-step sweeps (FP32 147.0-147.5 GLUPS at 64, 147.6-149.0 at 96;
// FP16 233.3-233.4 and 235.4-236.1).
static WaveGrid wave_grid_deep(const Layout &L, int cx, int storage, int levels) {
    const int64_t bands = (L.ly + WAVE_ROWS - 1) / WAVE_ROWS;
    const bool largest = bands * L.lx >= 128 * 6144;
    const int cmax = !largest ? 24 : levels >= 6 || storage != 0 ? 96 : 40;
    if (cx < 3 && largest && levels == 4 && storage == 0) cx = 63;
    // at least two bands of interior rows
    return make_thin(L, wave_grid_max(L, cx, cmax, 6144), levels - 1, 2 * WAVE_ROWS);
}

bool wave_grid_fits(const Layout &L) { return (int64_t)(3 + 16) * L.pitch * 8 < (int64_t)DROP; }

int wpanel_deep_cx(const Layout &L, int cx, int storage, int levels) {
    return wave_grid_deep(L, cx, storage, levels).cx;
}

bool wpanel_deep_possible(const Layout &L, int cx, int storage, int levels) {
    if (levels != 4 && levels != 6 && levels != 8) return false;
    if (!wave_grid_fits(L)) return false;
    WaveGrid g = wave_grid_deep(L, cx, storage, levels);
    // an interior of at least 124 rows and two chunks
    return g.thin && g.c_hi - g.c_lo >= 2;
}

template <typename S> constexpr const char *type_name() {
    return sizeof(S) == 8 ? "double" : sizeof(S) == 4 ? "float" : "_Float16";
}

// the interior kernel over g's interior waves, in workgroups of its WPG;
// a probe brackets it with events on st
template <typename S, int DEPTH, int LEVELS>
static void launch_in(const Problem &p, const WaveGrid &g, uint32_t dstep, const S *P, S *Q,
                      hipStream_t st, KernelProbe *probe) {
    constexpr int WPG = wpg_of<S, DEPTH, LEVELS>();
    if (probe) {
        std::snprintf(probe->name, sizeof probe->name, "k_wpanel_in<%s, %d, %d>", type_name<S>(),
                      DEPTH, LEVELS);
        // rows [r0, r1) of the columns the interior chunks [c_lo, c_hi) cover
        // (an aligned grid's leftmost chunk overlaps its neighbour)
        const int cols = in_chunk_x1(g, g.c_hi - 1) - (in_chunk_x1(g, g.c_lo) - g.cx);
        probe->nodes = (double)(g.r1 - g.r0) * (double)cols;
        (void)hipEventRecord(probe->t0, st);
    }
    k_wpanel_in<S, DEPTH, LEVELS><<<(g.nin + WPG - 1) / WPG, 64 * WPG, 0, st>>>(p, g, dstep, P, Q);
    if (probe) (void)hipEventRecord(probe->t1, st);
}

// Perimeter waves on `side`, interior waves on `st`, joined back into `st`:
// the two kernels write disjoint nodes and run concurrently.
hipError_t launch_wpanel(int storage, const Problem &p, const WaveGrid &g, const void *P, void *Q,
                         hipStream_t st, hipStream_t side, hipEvent_t fork, hipEvent_t join,
                         KernelProbe *probe) {
    if (g.cx < 3) return hipErrorInvalidValue;
    // two columns of loads in flight when the chunk width allows (three or
    // four measured no faster, for any storage format)
    const int depth = g.cx % 2 == 0 ? 2 : 1;
    hipError_t e = hipEventRecord(fork, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(side, fork, 0);
    if (e != hipSuccess) return e;
    auto go = [&](auto tag) {
        using S = decltype(tag);
        if (g.nedge && g.gl == 4)
            k_wpanel_edge<S, 4><<<g.nedge, 64, 0, side>>>(p, g, 0u, (const S *)P, (S *)Q);
        else if (g.nedge)
            k_wpanel_edge<S, 8><<<g.nedge, 64, 0, side>>>(p, g, 0u, (const S *)P, (S *)Q);
        if (!g.nin) return;
        const S *src = (const S *)P;
        S *dst = (S *)Q;
        if (depth == 2) launch_in<S, 2, 2>(p, g, 0u, src, dst, st, probe);
        else launch_in<S, 1, 2>(p, g, 0u, src, dst, st, probe);
    };
    if (storage == 0) go(double{});
    else if (storage == 1) go(float{});
    else go(_Float16{});
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(join, side);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, join, 0);
    return e;
}

// ---- the frame of the deep marches ---------------------------------------------
// The interior waves of a deep launch cover rows [H + 2, ly - H) and columns
// [H + 2, lx - H) (H = LEVELS - 1: their halo stays off the walls), in
// cx-column chunks right-aligned at lx - H, the leftmost moved right to start
// at H + 2 (it overlaps its neighbour, which writes the same values).  The
// frame around them advances two steps per pass (LEVELS / 2 passes through
// LEVELS / 2 - 1 further buffers); each pass but the last is widened inwards
// by the depth its successor reads (DeepFrame), so every pass reads only what
// the one before it wrote and no other kernel is needed.

// lanes per strip chunk of a frame pass with sr strip rows (and two halo rows)
constexpr int frame_gl(int sr) { return sr + 2 <= 8 ? 8 : 16; }

// columns per strip chunk of a deep march's frame: batched channels, and
// single lattices whose interior chunks are wider than 64 columns (the
// strips' two-step march runs a chunk's columns in sequence, so strips as
// wide as 96-column interior chunks made the three frame passes outlast the
// interior kernel).  Up to 64 columns the strips take the interior's width:
// 16-column strips cost the FP64 four-step sweep at 63 columns 2.5% (four
// times the perimeter waves beside its three-wave-per-SIMD interior: 122.4 /
// 123.2 -> 119.6 / 120.2 GLUPS), FP32 at 40 columns 0.7%.
constexpr int FRAME_CX = 16;
int deep_strip_cx(int cx) { return cx > 64 ? FRAME_CX : cx; }

// strip rows of pass j of a LEVELS-step frame (0 past the last pass)
constexpr int frame_rows(int levels, int j) {
    const int nmid = levels / 2 - 1;
    return j > nmid ? 0 : levels + (j < nmid ? ring_depth(nmid, j) : 0);
}

// Perimeter wave e of a pass of a batched deep march's frame (DeepFrame)
// with sr strip rows, nsw waves per strip and nsb side bands: the bottom
// strip's waves, the top strip's, then the left side's bands and the right
// side's.  GL: lanes per strip chunk (sr + 2 <= GL).
template <typename S, int GL>
__device__ __forceinline__ void frame_wave(const Problem &p, const WaveGrid &sg, int sr, int nsw,
                                           int nsb, int e, uint32_t dstep,
                                           const S *__restrict__ P, S *__restrict__ Q) {
    const Layout &L = p.lay;
    if (e < 2 * nsw) {
        const bool top = e >= nsw;
        const int ck0 = (top ? e - nsw : e) * (64 / GL);
        edge_march<S, GL>(p, sg, dstep, P, Q, top ? L.ly + 1 - sr : 1, sr, ck0);
        return;
    }
    e -= 2 * nsw;
    const bool right = e >= nsb;
    const int Y0 = 1 + sr + (right ? e - nsb : e) * WAVE_ROWS;
    const int Y1 = L.ly + 1 - sr;
    edge_march<S, 64>(p, sg, dstep, P, Q, Y0, Y1 - Y0 < WAVE_ROWS ? Y1 - Y0 : WAVE_ROWS, 0,
                      right ? L.lx + 1 - sr : 1, right ? L.lx + 1 : 1 + sr);
}

template <typename S, int GL>
__global__ __launch_bounds__(64) void k_frame(Problem p, WaveGrid sg, int sr, int nsw, int nsb,
                                              uint32_t dstep, const S *__restrict__ P,
                                              S *__restrict__ Q) {
    frame_wave<S, GL>(p, sg, sr, nsw, nsb, (int)blockIdx.x, dstep, P, Q);
}

// The interior grid and frame of a deep march of LEVELS = 4, 6 or 8 steps,
// interior chunks of cx columns, strips in chunks of strip_cx; false when the
// lattice has less than one chunk of interior columns or one interior row,
// or its widest strips or sides meet.
bool aligned_deep_grid(const Layout &L, int cx, int strip_cx, int levels, WaveGrid *g,
                       DeepFrame *f) {
    if ((levels != 4 && levels != 6 && levels != 8) || cx < 3 || !wave_grid_fits(L)) return false;
    const int H = levels - 1, nmid = levels / 2 - 1;
    const int xl = H + 2, xr = L.lx - H, r0 = H + 2, r1 = L.ly - H;
    const int sr0 = frame_rows(levels, 0);
    if (xr - xl < cx || r1 - r0 < 1 || 2 * sr0 > L.ly || 2 * sr0 > L.lx) return false;
    // 32-bit offsets of the interior kernel's buffer resources
    if ((int64_t)(cx + 16) * L.pitch * 8 >= (int64_t)DROP) return false;
    WaveGrid w = wave_grid_max(L, cx, cx, 1);
    f->sg = wave_grid_max(L, strip_cx, strip_cx, 1);
    f->nmid = nmid;
    for (int j = 0; j < 4; ++j) {
        const int sr = frame_rows(levels, j);
        f->sr[j] = sr;
        f->nsw[j] = j <= nmid ? (f->sg.ncx + 64 / frame_gl(sr) - 1) / (64 / frame_gl(sr)) : 0;
        f->nsb[j] = j <= nmid ? (L.ly - 2 * sr + WAVE_ROWS - 1) / WAVE_ROWS : 0;
    }
    w.xl = xl;
    w.xr = xr;
    w.ncx = (xr - xl + cx - 1) / cx;
    w.c_lo = 0;
    w.c_hi = w.ncx;
    w.r0 = r0;
    w.r1 = r1;
    w.nbl = (r1 - r0 + 64 - 2 * H - 1) / (64 - 2 * H);
    w.nin = w.nbl * w.ncx;
    w.b_lo = w.b_hi = 0;
    w.thin = 1;
    w.nthin = w.nedge = 0;
    w.gl = 0;
    *g = w;
    return true;
}

// perimeter waves of pass j
int batch_frame_waves(const DeepFrame &f, int j) { return 2 * f.nsw[j] + 2 * f.nsb[j]; }

hipError_t launch_wpanel_deep(int storage, const Problem &p, int cx, int levels, const void *P,
                              void *const *mid, void *Q, hipStream_t st, hipStream_t side,
                              hipEvent_t fork, hipEvent_t join, KernelProbe *probe) {
    if (!wpanel_deep_possible(p.lay, cx, storage, levels)) return hipErrorInvalidValue;
    const Layout &L = p.lay;
    // the chunk width of the thin-frame grid (its heuristics), the aligned grid
    // and frame over it, the strips in chunks of deep_strip_cx columns
    const int ccx = wpanel_deep_cx(L, cx, storage, levels);
    WaveGrid g;
    DeepFrame f;
    if (!aligned_deep_grid(L, ccx, deep_strip_cx(ccx), levels, &g, &f)) return hipErrorInvalidValue;
    const int nmid = levels / 2 - 1;
    // two columns of loads in flight when the chunk width allows, except for
    // the eight-step FP64 march (two columns in flight need 265 registers,
    // one wave per SIMD; one column 252, two waves)
    const int depth = g.cx % 2 == 0 && !(levels == 8 && storage == 0) ? 2 : 1;
    hipError_t e = hipEventRecord(fork, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(side, fork, 0);
    if (e != hipSuccess) return e;
    auto go = [&](auto tag) {
        using T = decltype(tag);
        // perimeter pass j advances the frame by steps 2j+1, 2j+2, from P or
        // mid[j-1] into mid[j] or Q
        for (int j = 0; j <= nmid; ++j) {
            const T *src = (const T *)(j == 0 ? P : mid[j - 1]);
            T *dst = (T *)(j == nmid ? Q : mid[j]);
            const int n = batch_frame_waves(f, j);
            if (frame_gl(f.sr[j]) == 8)
                k_frame<T, 8><<<n, 64, 0, side>>>(p, f.sg, f.sr[j], f.nsw[j], f.nsb[j], 2u * j,
                                                  src, dst);
            else
                k_frame<T, 16><<<n, 64, 0, side>>>(p, f.sg, f.sr[j], f.nsw[j], f.nsb[j], 2u * j,
                                                   src, dst);
        }
        const T *src = (const T *)P;
        T *dst = (T *)Q;
        if (levels == 4) {
            if (depth == 2) launch_in<T, 2, 4>(p, g, 0u, src, dst, st, probe);
            else launch_in<T, 1, 4>(p, g, 0u, src, dst, st, probe);
        } else if (levels == 6) {
            if (depth == 2) launch_in<T, 2, 6>(p, g, 0u, src, dst, st, probe);
            else launch_in<T, 1, 6>(p, g, 0u, src, dst, st, probe);
        } else {
            if (depth == 2) launch_in<T, 2, 8>(p, g, 0u, src, dst, st, probe);
            else launch_in<T, 1, 8>(p, g, 0u, src, dst, st, probe);
        }
    };
    if (storage == 0) go(double{});
    else if (storage == 1) go(float{});
    else go(_Float16{});
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(join, side);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, join, 0);
    return e;
}

// ---- many channels per launch ----------------------------------------------------
// The batched forms of the register marches: every channel of a batch has its
// own wave grid (ChannelWaves); one launch runs a pass of every channel's
// perimeter waves, one every channel's interior waves, from flat work
// tables.  An interior workgroup's waves belong to one
// channel (their lockstep barriers need the same march), and the workgroups
// of a channel are contiguous in the table, so xcd_block keeps a channel's
// neighbouring bands in one L2.  Each wave runs the single-lattice sweep's
// own code, so a channel evolves bit for bit as it would alone.

template <typename S, int GL>
__global__ __launch_bounds__(64) void k_batch_wedge(const ChannelWaves *__restrict__ ch,
                                                    const int2 *__restrict__ items, uint32_t dstep,
                                                    const S *__restrict__ A, S *__restrict__ B) {
    const int2 it = items[blockIdx.x];
    const ChannelWaves c = ch[__builtin_amdgcn_readfirstlane(it.x)];
    wpanel_edge_wave<S, GL>(c.p, c.g, __builtin_amdgcn_readfirstlane(it.y), dstep, A + c.off,
                            B + c.off);
}

template <typename S, int DEPTH, int LEVELS>
__global__ __launch_bounds__((64 * wpg_of<S, DEPTH, LEVELS>())) void k_batch_win(
    const ChannelWaves *__restrict__ ch, const int2 *__restrict__ items, uint32_t dstep,
    const S *__restrict__ A, S *__restrict__ B) {
    constexpr int WPG = wpg_of<S, DEPTH, LEVELS>();
    const int2 it = items[xcd_block(blockIdx.x, gridDim.x)];
    const ChannelWaves c = ch[__builtin_amdgcn_readfirstlane(it.x)];
    const int t = __builtin_amdgcn_readfirstlane(it.y * WPG + (int)(threadIdx.x >> 6));
    wpanel_in_wave<S, DEPTH, LEVELS>(c.p, c.g, t, dstep, A + c.off, B + c.off);
}

// Pass j's per-channel counts are read from the table in global memory: a
// private copy of the record indexed by the run-time j would live in
// scratch memory.
template <typename S, int GL>
__global__ __launch_bounds__(64) void k_batch_frame(const ChannelWaves *__restrict__ ch,
                                                    const int2 *__restrict__ items, int j,
                                                    uint32_t dstep, const S *__restrict__ A,
                                                    S *__restrict__ B) {
    const int2 it = items[blockIdx.x];
    const int ci = __builtin_amdgcn_readfirstlane(it.x);
    const ChannelWaves *cw = ch + ci;
    const int sr = cw->f.sr[j], nsw = cw->f.nsw[j], nsb = cw->f.nsb[j];
    const Problem p = cw->p;
    const WaveGrid sg = cw->f.sg;
    const int64_t off = cw->off;
    frame_wave<S, GL>(p, sg, sr, nsw, nsb, __builtin_amdgcn_readfirstlane(it.y), dstep, A + off,
                      B + off);
}

int batch_wpanel_wpg(int storage) {
    return storage == 0 ? wpg_of<double, 2, 2>() : storage == 1 ? wpg_of<float, 2, 2>()
                                                                : wpg_of<_Float16, 2, 2>();
}

// the batched deep marches run two columns of loads in flight (even chunk
// widths only)
int batch_deep_wpg(int storage, int levels) {
    if (levels == 4)
        return storage == 0 ? wpg_of<double, 2, 4>() : storage == 1 ? wpg_of<float, 2, 4>()
                                                                    : wpg_of<_Float16, 2, 4>();
    return storage == 0 ? wpg_of<double, 2, 6>() : storage == 1 ? wpg_of<float, 2, 6>()
                                                                : wpg_of<_Float16, 2, 6>();
}

// The interior grid and frame of a batched deep march: LEVELS = 4 or 6, even
// chunk widths (two columns of loads in flight), strips in FRAME_CX-column
// chunks whatever the interior's width: shorter marches, more waves (512
// channels, six steps in 64-column chunks: FP64 91.9 -> 97.1 GLUPS with
// 16-column strip chunks, 96.8 with 8; FP32 99.3 -> 100.7).
bool batch_deep_grid(const Layout &L, int cx, int levels, WaveGrid *g, DeepFrame *f) {
    if ((levels != 4 && levels != 6) || cx < 4 || cx % 2) return false;
    return aligned_deep_grid(L, cx, FRAME_CX, levels, g, f);
}

// one two-step pass of the tables' channels (perimeter then interior waves,
// both on `q`: they write disjoint nodes)
template <typename S>
static void batch_pair_on(const BatchTables &t, uint32_t dstep, const S *A, S *B, hipStream_t q) {
    if (t.nedge > 0) k_batch_wedge<S, 4><<<t.nedge, 64, 0, q>>>(t.ch, t.edge, dstep, A, B);
    if (t.nin_wgs > 0)
        k_batch_win<S, 2, 2><<<t.nin_wgs, 64 * wpg_of<S, 2, 2>(), 0, q>>>(t.ch, t.in, dstep, A, B);
}

// perimeter waves on `side`, interior waves on `st`, joined back into `st`
// (they write disjoint nodes), as the single-lattice panel launch
hipError_t launch_batch_wpanel(int storage, const BatchTables &t, uint32_t dstep, const void *A,
                               void *B, hipStream_t st, hipStream_t side, hipEvent_t fork,
                               hipEvent_t join) {
    hipError_t e = hipEventRecord(fork, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(side, fork, 0);
    if (e != hipSuccess) return e;
    auto go = [&](auto tag) {
        using S = decltype(tag);
        if (t.nedge > 0)
            k_batch_wedge<S, 4><<<t.nedge, 64, 0, side>>>(t.ch, t.edge, dstep, (const S *)A, (S *)B);
        if (t.nin_wgs > 0)
            k_batch_win<S, 2, 2><<<t.nin_wgs, 64 * wpg_of<S, 2, 2>(), 0, st>>>(
                t.ch, t.in, dstep, (const S *)A, (S *)B);
    };
    if (storage == 0) go(double{});
    else if (storage == 1) go(float{});
    else go(_Float16{});
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(join, side);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, join, 0);
    return e;
}

// The deep channels' interior waves march LEVELS steps on `st` (A -> B);
// on `side`, their frame advances two steps per pass (A -> mid[0] -> ... ->
// B, each pass but the last widened by what the next one reads), then the
// shallow channels take LEVELS/2 two-step passes through the same buffers.
hipError_t launch_batch_wpanel_deep(int storage, int levels, const BatchTables &deep,
                                    const int2 *const *frame_items, const int *nframe,
                                    const BatchTables &shallow, const void *A, void *const *mid,
                                    void *B, hipStream_t st, hipStream_t side, hipEvent_t fork,
                                    hipEvent_t join) {
    if (levels != 4 && levels != 6) return hipErrorInvalidValue;
    const int nmid = levels / 2 - 1;
    hipError_t e = hipEventRecord(fork, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(side, fork, 0);
    if (e != hipSuccess) return e;
    auto go = [&](auto tag) {
        using S = decltype(tag);
        auto src_of = [&](int j) { return (const S *)(j == 0 ? A : mid[j - 1]); };
        auto dst_of = [&](int j) { return (S *)(j == nmid ? B : mid[j]); };
        // the interior waves are enqueued first: queued behind the frame's
        // perimeter waves they waited for slots while those ran at low
        // occupancy (512 channels, FP32 four steps: 91.7 -> 98.5 GLUPS,
        // FP64 six steps 85.3 -> 87.5, FP16 four steps 127.2 -> 134.8)
        if (deep.nin_wgs > 0) {
            if (levels == 4)
                k_batch_win<S, 2, 4><<<deep.nin_wgs, 64 * wpg_of<S, 2, 4>(), 0, st>>>(
                    deep.ch, deep.in, 0u, (const S *)A, (S *)B);
            else
                k_batch_win<S, 2, 6><<<deep.nin_wgs, 64 * wpg_of<S, 2, 6>(), 0, st>>>(
                    deep.ch, deep.in, 0u, (const S *)A, (S *)B);
        }
        for (int j = 0; j <= nmid; ++j) {
            if (nframe[j] <= 0) continue;
            if (frame_gl(frame_rows(levels, j)) == 8)
                k_batch_frame<S, 8><<<nframe[j], 64, 0, side>>>(deep.ch, frame_items[j], j, 2u * j,
                                                               src_of(j), dst_of(j));
            else
                k_batch_frame<S, 16><<<nframe[j], 64, 0, side>>>(deep.ch, frame_items[j], j,
                                                                2u * j, src_of(j), dst_of(j));
        }
        for (int j = 0; j <= nmid; ++j) batch_pair_on<S>(shallow, 2u * j, src_of(j), dst_of(j), side);
    };
    if (storage == 0) go(double{});
    else if (storage == 1) go(float{});
    else go(_Float16{});
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(join, side);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, join, 0);
    return e;
}

}  // namespace lbm
