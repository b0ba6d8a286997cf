// Host-only check of the deep marches' geometry (batch_deep_grid and
// aligned_deep_grid, batch_frame_waves, and the perimeter waves' row and
// column ranges, restated here from frame_wave / edge_march), for many
// lattice shapes, chunk widths and depths, batched and single-lattice forms:
//  * every pass's perimeter waves own disjoint nodes;
//  * the last pass's frame and the interior waves' owned nodes tile the
//    channel exactly (interior chunks may overlap each other: they write the
//    same values);
//  * pass j reads (its nodes and two more in every direction) only nodes that
//    pass j - 1 wrote;
//  * interior waves read only columns 0 .. lx + 1 and rows whose halo stays
//    off the walls;
//  * every lattice the single-lattice deep sweeps accept gets such a grid at
//    the chunk width its launch takes.
// Test infrastructure: built and run by tests/test_frame_geometry.py; no GPU.
#include <algorithm>
#include <cstdio>
#include <vector>
#include <cstdlib>
#include "launch.hpp"
using namespace lbm;
namespace lbm { Layout make_layout(int lx, int ly); }

// strip_cx 0: the batched form (batch_deep_grid); else the single-lattice
// form with strips in strip_cx-column chunks
static int check(int lx, int ly, int cx, int levels, int strip_cx) {
    Layout L = make_layout(lx, ly);
    WaveGrid g; DeepFrame f;
    if (strip_cx ? !aligned_deep_grid(L, cx, strip_cx, levels, &g, &f)
                 : !batch_deep_grid(L, cx, levels, &g, &f))
        return -1;
    const int H = levels - 1, nmid = levels / 2 - 1;
    auto idx = [&](int x, int y) { return (size_t)x * (ly + 2) + y; };
    std::vector<std::vector<int>> reg(nmid + 1, std::vector<int>((size_t)(lx + 2) * (ly + 2), 0));
    for (int j = 0; j <= nmid; ++j) {
        const int sr = f.sr[j];
        const int GL = sr + 2 <= 8 ? 8 : 16;
        if (sr + 2 > 16) { printf("strip of %d rows beyond 16 lanes\n", sr); return 1; }
        const int nw = batch_frame_waves(f, j);
        for (int e = 0; e < nw; ++e) {
            int ee = e;
            if (ee < 2 * f.nsw[j]) {
                bool top = ee >= f.nsw[j];
                int ck0 = (top ? ee - f.nsw[j] : ee) * (64 / GL);
                int Y0 = top ? ly + 1 - sr : 1, nrows = sr;
                if (nrows + 2 > GL) { printf("GL too small\n"); return 1; }
                for (int lane = 0; lane < 64; ++lane) {
                    int lr = lane & (GL - 1), ckl = ck0 + lane / GL;
                    if (ckl >= f.sg.ncx) continue;
                    WaveSpan w = wave_span(L, f.sg, 0, ckl);
                    int y = Y0 - 1 + lr;
                    bool owner = lr >= 1 && lr <= nrows && y <= ly;
                    if (!owner) continue;
                    for (int x = w.X0; x < w.X1; ++x) reg[j][idx(x, y)]++;
                }
            } else {
                ee -= 2 * f.nsw[j];
                bool right = ee >= f.nsb[j];
                int Y0 = 1 + sr + (right ? ee - f.nsb[j] : ee) * WAVE_ROWS;
                int Y1 = ly + 1 - sr;
                int nrows = Y1 - Y0 < WAVE_ROWS ? Y1 - Y0 : WAVE_ROWS;
                int xa = right ? lx + 1 - sr : 1, xb = right ? lx + 1 : 1 + sr;
                if (nrows <= 0) { printf("empty side band %d %d j%d\n", lx, ly, j); return 1; }
                for (int y = Y0; y < Y0 + nrows; ++y)
                    for (int x = xa; x < xb; ++x) reg[j][idx(x, y)]++;
            }
        }
        for (int x = 1; x <= lx; ++x)
            for (int y = 1; y <= ly; ++y)
                if (reg[j][idx(x, y)] > 1) { printf("frame overlap %dx%d j%d at %d,%d\n", lx, ly, j, x, y); return 1; }
    }
    // interior (final level)
    std::vector<int> in((size_t)(lx + 2) * (ly + 2), 0);
    const int ROWS = 64 - 2 * H;
    for (int t = 0; t < g.nin; ++t) {
        int ck = g.c_lo + t / g.nbl;
        int Y0 = g.r0 + (t % g.nbl) * ROWS;
        int X1 = in_chunk_x1(g, ck), X0 = X1 - g.cx;
        if (X0 - H - 1 < 1 || X1 + H > lx) { printf("interior reach %dx%d ck%d X0 %d X1 %d\n", lx, ly, ck, X0, X1); return 1; }
        if (X0 < H + 2 || X1 > lx - H) { printf("interior chunk off %d %d\n", X0, X1); return 1; }
        for (int lane = H; lane < H + ROWS; ++lane) {
            int y = Y0 - H + lane;
            if (y >= g.r1) continue;
            if (y - H < 2 || y + H > ly - 1) { printf("interior row reach %d\n", y); return 1; }
            for (int x = X0; x < X1; ++x) in[idx(x, y)] = 1;
        }
    }
    for (int x = 1; x <= lx; ++x)
        for (int y = 1; y <= ly; ++y) {
            int a = reg[nmid][idx(x, y)], b = in[idx(x, y)];
            if (a + b != 1) { printf("final cover %dx%d at %d,%d frame %d in %d\n", lx, ly, x, y, a, b); return 1; }
        }
    // pass j needs pass j-1 output at region(j) dilated by 2
    for (int j = 1; j <= nmid; ++j)
        for (int x = 1; x <= lx; ++x)
            for (int y = 1; y <= ly; ++y) {
                if (!reg[j][idx(x, y)]) continue;
                for (int dx = -2; dx <= 2; ++dx)
                    for (int dy = -2; dy <= 2; ++dy) {
                        int u = x + dx, v = y + dy;
                        if (u < 1 || u > lx || v < 1 || v > ly) continue;
                        if (!reg[j - 1][idx(u, v)]) { printf("reach %dx%d j%d (%d,%d) needs (%d,%d)\n", lx, ly, j, x, y, u, v); return 1; }
                    }
            }
    return 0;
}

// every lattice the single-lattice deep sweeps accept (wpanel_deep_possible)
// gets an aligned grid at the chunk width the launch takes
static int check_eligible(int lx, int ly, int storage, int levels) {
    Layout L = make_layout(lx, ly);
    if (!wpanel_deep_possible(L, 0, storage, levels)) return -1;
    const int cx = wpanel_deep_cx(L, 0, storage, levels);
    WaveGrid g; DeepFrame f;
    if (!aligned_deep_grid(L, cx, deep_strip_cx(cx), levels, &g, &f)) {
        printf("eligible %dx%d storage %d levels %d but no aligned grid (cx %d)\n", lx, ly, storage, levels, cx);
        return 1;
    }
    return check(lx, ly, cx, levels, deep_strip_cx(cx));
}

int main() {
    int bad = 0, ok = 0, skip = 0;
    srand(1);
    std::vector<std::pair<int,int>> dims = {{700,250},{1030,140},{520,300},{130,33},{300,75},{64,16},{48,300},{17,5},{9,4},{128,32},{2048,512},{36,20},{40,20},{200,22}};
    for (int k = 0; k < 80; ++k) { int lx = 20 + rand() % 2100; dims.push_back({lx, std::max(3, lx / 4 + rand() % 7 - 3)}); }
    for (auto [lx, ly] : dims) {
        for (int levels : {4, 6})
            for (int cx : {4, 6, 8, 12, 16, 24, 32, 48, 64}) {
                int r = check(lx, ly, cx, levels, 0);
                if (r < 0) ++skip; else if (r) { ++bad; printf("  bad: %dx%d cx %d levels %d\n", lx, ly, cx, levels); } else ++ok;
            }
        for (int levels : {4, 6, 8})
            for (int cx : {3, 7, 16, 24, 39, 63, 64}) {
                int r = check(lx, ly, cx, levels, cx);
                if (r < 0) ++skip; else if (r) { ++bad; printf("  bad: %dx%d cx %d levels %d (strips of cx)\n", lx, ly, cx, levels); } else ++ok;
            }
    }
    // the deep sweeps' eligibility on lattice shapes around its thresholds
    for (int ly : {120, 130, 135, 136, 137, 140, 144, 150, 200, 256, 333, 700})
        for (int lx : {40, 64, 100, 128, 200, 300, 512, 700, 1500})
            for (int storage : {0, 1, 2})
                for (int levels : {4, 6, 8}) {
                    int r = check_eligible(lx, ly, storage, levels);
                    if (r < 0) ++skip; else if (r) { ++bad; printf("  bad: eligible %dx%d\n", lx, ly); } else ++ok;
                }
    // the largest lattice at the chunk widths its deep launches take or may be
    // given, strips as wide as the launch makes them
    for (int levels : {4, 6, 8}) {
        int r = check_eligible(16384, 4096, 0, levels);
        if (r < 0) ++skip; else if (r) { ++bad; printf("  bad: eligible 16384x4096 levels %d\n", levels); } else ++ok;
        for (int cx : {40, 63, 64, 96, 128}) {
            r = check(16384, 4096, cx, levels, deep_strip_cx(cx));
            if (r < 0) ++skip; else if (r) { ++bad; printf("  bad: 16384x4096 cx %d levels %d\n", cx, levels); } else ++ok;
        }
    }
    printf("ok %d bad %d skipped %d\n", ok, bad, skip);
    return bad != 0;
}
