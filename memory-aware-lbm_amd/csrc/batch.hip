// batch.hip -- many independent channels advanced by one persistent launch
// per time step.
//
// All channels live in two pools (ping-pong), each channel a self-contained
// SoA lattice with its own ghost ring at an aligned offset.  The work of one
// step is a flat table of tiles: tile = (channel, ty rows x tx columns) with
// tx * ty = 256 nodes, ty a power of two >= 32 chosen per channel so short
// channels do not leave most lanes idle.  A launch has nwg = CUs x
// wgs_per_cu workgroups; workgroup b walks the contiguous item range
// [b * per_wg, (b+1) * per_wg), so neighbouring tiles of a channel stay on one
// XCD's L2.  Each node runs exactly the single-lattice node step
// (node_step.hpp), so every channel's result is bit-identical to running it
// alone.  K consecutive steps can be captured into a hipGraph and replayed.
//
// Batches can instead advance two steps per launch (depth 2, the default):
// one workgroup per two-step LDS tile of the single-lattice tight_block sweep
// (block_tile.hpp), the tiles of all channels in one flat table, renumbered
// so each XCD takes a contiguous range.  The same tile body, so the same
// bits; the lattice crosses HBM once per two steps.
//
// No counterpart in the reference (one lattice per process there).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/lbm_hip.h"
#include "host_util.hpp"
#include "block_tile.hpp"
#include "launch.hpp"
#include "node_step.hpp"

namespace lbm {

struct BatchEntry {
    Problem p;
    int64_t off;    // element offset of the channel in each pool
    int64_t item0;  // first work item of the channel
    int ty, tx, nty;
    // two-step tiles: btx columns x (BATCH_TYP - 2) rows, bnbx x bnby of
    // them, numbered from tile0 (column strip major)
    int64_t tile0;
    int btx, bnbx, bnby;
};

constexpr int BATCH_TYP = 64;  // two-step tile rows + the one-node ring

// LDS of one two-step tile: compute-type values (8 bytes for FP64 storage,
// 4 for FP32 / FP16)
inline size_t batch_block_lds(int tile_cols, int storage) {
    return 9 * (size_t)(tile_cols + 2) * BATCH_TYP * (storage == LBM_STORE_F64 ? 8 : 4);
}

template <typename S, bool PULL>
__global__ __launch_bounds__(256) void k_batch(const BatchEntry *__restrict__ ents,
                                               const int *__restrict__ item_chan, int64_t nitems,
                                               int64_t per_wg, const S *__restrict__ A,
                                               S *__restrict__ B, const uint32_t *__restrict__ stepctr) {
    const uint32_t step = sizeof(S) == 4 ? *stepctr : 0u;
    int64_t it = (int64_t)blockIdx.x * per_wg;
    int64_t end = it + per_wg < nitems ? it + per_wg : nitems;
    for (; it < end; ++it) {
        const int c = __builtin_amdgcn_readfirstlane(item_chan[it]);
        const BatchEntry &e = ents[c];
        const int ty = e.ty, nty = e.nty;
        int64_t local = it - e.item0;
        int xt = (int)(local / nty);
        int yt = (int)(local - (int64_t)xt * nty);
        int x = 1 + xt * e.tx + (int)threadIdx.x / ty;
        int y = 1 + yt * ty + ((int)threadIdx.x & (ty - 1));
        if (x > e.p.lay.lx || y > e.p.lay.ly) continue;
        if constexpr (sizeof(S) == 8)
            node_step_f64<PULL>(e.p, x, y, (const double *)A + e.off, (double *)B + e.off);
        else
            node_step_lp<S, PULL, false>(e.p, x, y, A + e.off, B + e.off, step);
    }
}

// Two steps of every channel, one workgroup per tile (collided state in A,
// collided state two steps later in B).
template <typename S>
__global__ __launch_bounds__(BLOCK_THREADS) void k_batch_block(const BatchEntry *__restrict__ ents,
                                                             const int *__restrict__ tile_chan,
                                                             const S *__restrict__ A,
                                                             S *__restrict__ B) {
    extern __shared__ __align__(16) unsigned char lds_raw[];  // [9][btx+2][BATCH_TYP]
    const int t = xcd_block(blockIdx.x, gridDim.x);
    const int c = __builtin_amdgcn_readfirstlane(tile_chan[t]);
    const BatchEntry &e = ents[c];
    const int local = t - (int)e.tile0;
    int by = local % e.bnby, bx = local / e.bnby;
    if constexpr (sizeof(S) == 4) {
        pin_index(by);
        pin_index(bx);
    }
    const uint32_t m = sizeof(S) == 4 ? *e.p.step : 0u;   // the launch's first collision
    block_tile<S, BATCH_TYP>(e.p, e.btx, e.bnbx, bx, by, m, A + e.off, B + e.off,
                             (typename Calc<S>::T *)lds_raw);
}

template <typename S>
static hipError_t batch_launch(bool pull, int nwg, const BatchEntry *ents, const int *item_chan,
                               int64_t nitems, const void *A, void *B, const uint32_t *step,
                               hipStream_t st) {
    int64_t per_wg = (nitems + nwg - 1) / nwg;
    if (pull)
        k_batch<S, true><<<nwg, 256, 0, st>>>(ents, item_chan, nitems, per_wg, (const S *)A,
                                             (S *)B, step);
    else
        k_batch<S, false><<<nwg, 256, 0, st>>>(ents, item_chan, nitems, per_wg, (const S *)A,
                                              (S *)B, step);
    return hipGetLastError();
}

}  // namespace lbm

using namespace lbm;

namespace {
size_t esize(int storage) { return storage == LBM_STORE_F64 ? 8 : storage == LBM_STORE_F32 ? 4 : 2; }
}

struct lbm_batch {
    int n;
    int storage;
    std::vector<BatchEntry> ents;
    BatchEntry *d_ents = nullptr;
    int *d_item_chan = nullptr;
    int64_t nitems = 0;
    void *pool[2] = {nullptr, nullptr};
    double *params = nullptr;
    int64_t pool_elems = 0;
    int64_t cells = 0;
    int cur = 0;
    int collided = 0;
    int64_t steps = 0;
    uint32_t *d_step = nullptr;   // collision counter (keys FP32 stochastic rounding)
    hipStream_t stream = nullptr;
    int num_cu = 256;
    int wgs_per_cu = 8;
    int ty_cap = 256;
    int graph_steps = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // by starting buffer
    int graph_flip[2] = {0, 0};                     // buffer swaps per replay, mod 2
    int depth = 1;                                  // steps per launch (2: LDS tiles)
    int tile_cols = 8;                              // columns of a two-step tile (set
                                                    // per storage format at creation)
    int min_lx = 0;                                 // two-step tiles need lx >= 4
    int *d_tile_chan = nullptr;                     // channel of each two-step tile
    int64_t ntiles = 0;
    std::vector<int> ly_of, lx_of;
    // multi-step sweep: LDS tiles (LBM_TIGHT_BLOCK, two steps), the register
    // march (LBM_PANEL, two steps) with its per-channel wave grids and work
    // tables, or the deep register marches (LBM_PANEL4 / LBM_PANEL6, four /
    // six steps) with theirs beside the two-step ones (their channels without
    // a deep interior, and remainders of two steps)
    int sweep = LBM_TIGHT_BLOCK;
    int panel_cx = 6;                               // the two-step march's chunk width
    hipStream_t side = nullptr;                     // its perimeter waves' stream
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<ChannelWaves> cw;                   // two-step grids, every channel
    ChannelWaves *d_cw = nullptr;
    int2 *d_edge_items = nullptr, *d_in_items = nullptr;
    int nedge_items = 0, nin_wgs = 0;
    int deep_levels = 0;                            // 4 or 6 once a deep sweep was set
    int deep_cx = 24;                               // the deep marches' chunk width
    int ndeep = 0;                                  // channels with a deep grid
    ChannelWaves *d_dcw = nullptr;                  // their deep grids
    int2 *d_din = nullptr, *d_frame[3] = {nullptr, nullptr, nullptr};
    int ndin_wgs = 0, nframe[3] = {0, 0, 0};
    int2 *d_sedge = nullptr, *d_sin = nullptr;      // the other channels' two-step tables
    int nsedge = 0, nsin_wgs = 0;
    void *mid[2] = {nullptr, nullptr};              // the deep launches' further buffers
};

namespace {

int ceil_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

void drop_graphs(lbm_batch *b) {
    for (int k = 0; k < 2; ++k)
        if (b->graph[k]) {
            (void)hipGraphExecDestroy(b->graph[k]);
            b->graph[k] = nullptr;
        }
}

// tile shapes and the work tables for the current ty_cap and tile_cols
int build_items(lbm_batch *b) {
    std::vector<int> chan, tchan;
    int64_t item = 0, tile = 0;
    for (int c = 0; c < b->n; ++c) {
        BatchEntry &e = b->ents[c];
        int ty = std::min(b->ty_cap, std::max(32, ceil_pow2(e.p.lay.ly)));
        e.ty = ty;
        e.tx = 256 / ty;
        e.nty = (e.p.lay.ly + ty - 1) / ty;
        int ntx = (e.p.lay.lx + e.tx - 1) / e.tx;
        e.item0 = item;
        int64_t cnt = (int64_t)ntx * e.nty;
        chan.insert(chan.end(), (size_t)cnt, c);
        item += cnt;
        e.btx = b->tile_cols;
        e.bnbx = (e.p.lay.lx + e.btx - 1) / e.btx;
        e.bnby = (e.p.lay.ly + BATCH_TYP - 3) / (BATCH_TYP - 2);
        e.tile0 = tile;
        int64_t nt = (int64_t)e.bnbx * e.bnby;
        tchan.insert(tchan.end(), (size_t)nt, c);
        tile += nt;
    }
    if (tile > INT32_MAX) return fail(LBM_ERR_ARG, "batch has more than 2^31 two-step tiles");
    if (b->d_item_chan) (void)hipFree(b->d_item_chan);
    if (b->d_tile_chan) (void)hipFree(b->d_tile_chan);
    b->d_item_chan = nullptr;
    b->d_tile_chan = nullptr;
    b->nitems = item;
    b->ntiles = tile;
    HIPCHK(hipMalloc((void **)&b->d_item_chan, chan.size() * sizeof(int)));
    HIPCHK(hipMemcpy(b->d_item_chan, chan.data(), chan.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc((void **)&b->d_tile_chan, tchan.size() * sizeof(int)));
    HIPCHK(hipMemcpy(b->d_tile_chan, tchan.data(), tchan.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_ents, b->ents.data(), b->ents.size() * sizeof(BatchEntry),
                     hipMemcpyHostToDevice));
    drop_graphs(b);
    return LBM_OK;
}

// the register march's per-channel wave grids and work tables
int build_waves(lbm_batch *b) {
    std::vector<ChannelWaves> cw(b->n);
    std::vector<int2> edge, in;
    const int wpg = batch_wpanel_wpg(b->storage);
    for (int c = 0; c < b->n; ++c) {
        const BatchEntry &e = b->ents[c];
        if (e.p.lay.lx < 4 || !wave_grid_fits(e.p.lay))
            return fail(LBM_ERR_UNSUPPORTED, "channel " + std::to_string(c) +
                                                 " is too small or too tall for the register march");
        cw[c].p = e.p;
        cw[c].g = wave_grid(e.p.lay, b->panel_cx);
        cw[c].off = e.off;
        for (int k = 0; k < cw[c].g.nedge; ++k) edge.push_back(make_int2(c, k));
        for (int k = 0; k < (cw[c].g.nin + wpg - 1) / wpg; ++k) in.push_back(make_int2(c, k));
    }
    if (edge.size() > INT32_MAX || in.size() > INT32_MAX)
        return fail(LBM_ERR_ARG, "batch has more than 2^31 waves");
    if (b->d_cw) (void)hipFree(b->d_cw);
    if (b->d_edge_items) (void)hipFree(b->d_edge_items);
    if (b->d_in_items) (void)hipFree(b->d_in_items);
    b->d_cw = nullptr;
    b->d_edge_items = b->d_in_items = nullptr;
    HIPCHK(hipMalloc((void **)&b->d_cw, cw.size() * sizeof(ChannelWaves)));
    HIPCHK(hipMemcpy(b->d_cw, cw.data(), cw.size() * sizeof(ChannelWaves), hipMemcpyHostToDevice));
    if (!edge.empty()) {
        HIPCHK(hipMalloc((void **)&b->d_edge_items, edge.size() * sizeof(int2)));
        HIPCHK(hipMemcpy(b->d_edge_items, edge.data(), edge.size() * sizeof(int2),
                         hipMemcpyHostToDevice));
    }
    if (!in.empty()) {
        HIPCHK(hipMalloc((void **)&b->d_in_items, in.size() * sizeof(int2)));
        HIPCHK(hipMemcpy(b->d_in_items, in.data(), in.size() * sizeof(int2),
                         hipMemcpyHostToDevice));
    }
    b->nedge_items = (int)edge.size();
    b->nin_wgs = (int)in.size();
    b->cw = std::move(cw);
    return LBM_OK;
}

// v to a fresh device array *d (nullptr when empty), the old one freed
template <typename T> int upload(const std::vector<T> &v, T **d) {
    if (*d) (void)hipFree(*d);
    *d = nullptr;
    if (v.empty()) return LBM_OK;
    HIPCHK(hipMalloc((void **)d, v.size() * sizeof(T)));
    HIPCHK(hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return LBM_OK;
}

// the deep marches' grids and work tables (deep_levels steps, deep_cx-column
// chunks) for the channels that take them, the two-step tables (b->cw,
// built first) of the others, and the further buffers
int build_deep(lbm_batch *b) {
    const int levels = b->deep_levels, nmid = levels / 2 - 1;
    const int wpg = batch_deep_wpg(b->storage, levels), swpg = batch_wpanel_wpg(b->storage);
    std::vector<ChannelWaves> dcw;
    std::vector<int2> in, frame[3], sedge, sin;
    for (int c = 0; c < b->n; ++c) {
        const BatchEntry &e = b->ents[c];
        ChannelWaves w;
        std::memset(&w, 0, sizeof w);
        w.p = e.p;
        w.off = e.off;
        if (batch_deep_grid(e.p.lay, b->deep_cx, levels, &w.g, &w.f)) {
            const int d = (int)dcw.size();
            dcw.push_back(w);
            for (int k = 0; k < (w.g.nin + wpg - 1) / wpg; ++k) in.push_back(make_int2(d, k));
            for (int j = 0; j <= nmid; ++j)
                for (int k = 0; k < batch_frame_waves(w.f, j); ++k) frame[j].push_back(make_int2(d, k));
        } else {
            const WaveGrid &g = b->cw[c].g;
            for (int k = 0; k < g.nedge; ++k) sedge.push_back(make_int2(c, k));
            for (int k = 0; k < (g.nin + swpg - 1) / swpg; ++k) sin.push_back(make_int2(c, k));
        }
    }
    for (const auto *v : {&in, &frame[0], &frame[1], &frame[2], &sedge, &sin})
        if (v->size() > INT32_MAX) return fail(LBM_ERR_ARG, "batch has more than 2^31 waves");
    int rc = upload(dcw, &b->d_dcw);
    if (!rc) rc = upload(in, &b->d_din);
    for (int j = 0; j < 3 && !rc; ++j) rc = upload(frame[j], &b->d_frame[j]);
    if (!rc) rc = upload(sedge, &b->d_sedge);
    if (!rc) rc = upload(sin, &b->d_sin);
    for (int j = 0; j < nmid && !rc; ++j)
        if (!b->mid[j]) {
            hipError_t e = hipMalloc(&b->mid[j], (size_t)b->pool_elems * esize(b->storage));
            if (e != hipSuccess) { b->mid[j] = nullptr; rc = hip_fail(e, "hipMalloc deep buffer"); }
        }
    if (rc) {
        b->ndeep = b->ndin_wgs = b->nsedge = b->nsin_wgs = 0;
        b->nframe[0] = b->nframe[1] = b->nframe[2] = 0;
        return rc;
    }
    b->ndeep = (int)dcw.size();
    b->ndin_wgs = (int)in.size();
    for (int j = 0; j < 3; ++j) b->nframe[j] = (int)frame[j].size();
    b->nsedge = (int)sedge.size();
    b->nsin_wgs = (int)sin.size();
    return LBM_OK;
}

int sweep_levels(int sweep) { return sweep == LBM_PANEL4 ? 4 : sweep == LBM_PANEL6 ? 6 : 2; }

hipError_t launch_step(lbm_batch *b, bool pull) {
    int nwg = std::max(1, b->num_cu * b->wgs_per_cu);
    const void *src = b->pool[b->cur];
    void *dst = b->pool[1 - b->cur];
    hipError_t e;
    switch (b->storage) {
    case LBM_STORE_F64:
        return batch_launch<double>(pull, nwg, b->d_ents, b->d_item_chan, b->nitems, src, dst,
                                    b->d_step, b->stream);
    case LBM_STORE_F32:
        e = batch_launch<float>(pull, nwg, b->d_ents, b->d_item_chan, b->nitems, src, dst,
                                b->d_step, b->stream);
        break;
    default:
        e = batch_launch<_Float16>(pull, nwg, b->d_ents, b->d_item_chan, b->nitems, src, dst,
                                   b->d_step, b->stream);
        break;
    }
    return e == hipSuccess ? launch_bump(b->d_step, 1, b->stream) : e;
}

// two steps of every channel from the collided state
hipError_t launch_pair(lbm_batch *b) {
    const dim3 grid((unsigned)b->ntiles);
    const size_t lds = batch_block_lds(b->tile_cols, b->storage);
    const void *src = b->pool[b->cur];
    void *dst = b->pool[1 - b->cur];
    if (b->sweep != LBM_TIGHT_BLOCK) {
        const BatchTables t{b->d_cw, b->d_edge_items, b->nedge_items, b->d_in_items, b->nin_wgs};
        hipError_t e = launch_batch_wpanel(b->storage, t, 0u, src, dst, b->stream, b->side,
                                           b->fork, b->join);
        if (e == hipSuccess && b->storage != LBM_STORE_F64) e = launch_bump(b->d_step, 2, b->stream);
        return e;
    }
    switch (b->storage) {
    case LBM_STORE_F64:
        k_batch_block<double><<<grid, BLOCK_THREADS, lds, b->stream>>>(
            b->d_ents, b->d_tile_chan, (const double *)src, (double *)dst);
        return hipGetLastError();
    case LBM_STORE_F32:
        k_batch_block<float><<<grid, BLOCK_THREADS, lds, b->stream>>>(
            b->d_ents, b->d_tile_chan, (const float *)src, (float *)dst);
        break;
    default:
        k_batch_block<_Float16><<<grid, BLOCK_THREADS, lds, b->stream>>>(
            b->d_ents, b->d_tile_chan, (const _Float16 *)src, (_Float16 *)dst);
        break;
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? launch_bump(b->d_step, 2, b->stream) : e;
}

// four or six steps of every channel from the collided state
hipError_t launch_deep(lbm_batch *b) {
    const int levels = sweep_levels(b->sweep);
    const BatchTables deep{b->d_dcw, nullptr, 0, b->d_din, b->ndin_wgs};
    const BatchTables shallow{b->d_cw, b->d_sedge, b->nsedge, b->d_sin, b->nsin_wgs};
    const int2 *const frames[3] = {b->d_frame[0], b->d_frame[1], b->d_frame[2]};
    hipError_t e = launch_batch_wpanel_deep(b->storage, levels, deep, frames, b->nframe, shallow,
                                            b->pool[b->cur], b->mid, b->pool[1 - b->cur],
                                            b->stream, b->side, b->fork, b->join);
    if (e == hipSuccess && b->storage != LBM_STORE_F64)
        e = launch_bump(b->d_step, (uint32_t)levels, b->stream);
    return e;
}

// steps of the next launch: the sweep's depth where the setting and the
// state allow it, else two (the two-step form of the sweep) or one
int launch_depth(const lbm_batch *b, int left, bool collided) {
    if (b->depth != 2 || !collided || left < 2) return 1;
    const int levels = sweep_levels(b->sweep);
    return left >= levels ? levels : 2;
}

hipError_t launch_k(lbm_batch *b, int k, bool collided) {
    return k == 1 ? launch_step(b, collided) : k == 2 ? launch_pair(b) : launch_deep(b);
}

int eager_steps(lbm_batch *b, int nsteps) {
    while (nsteps > 0) {
        const int k = launch_depth(b, nsteps, b->collided != 0);
        HIPCHK(launch_k(b, k, b->collided != 0));
        b->cur = 1 - b->cur;
        b->collided = 1;
        b->steps += k;
        nsteps -= k;
    }
    return LBM_OK;
}

// capture graph_steps pulls starting from buffer `start`
int capture(lbm_batch *b, int start) {
    int saved_cur = b->cur;
    b->cur = start;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    hipError_t e = hipSuccess;
    int flips = 0;
    for (int left = b->graph_steps; left > 0 && e == hipSuccess;) {
        const int k = launch_depth(b, left, true);
        e = launch_k(b, k, true);
        b->cur = 1 - b->cur;
        ++flips;
        left -= k;
    }
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    b->cur = saved_cur;
    b->graph_flip[start] = flips & 1;
    if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return hip_fail(e, "capture launch"); }
    if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
    hipError_t e3 = hipGraphInstantiate(&b->graph[start], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e3 != hipSuccess) return hip_fail(e3, "hipGraphInstantiate");
    return LBM_OK;
}

int step_impl(lbm_batch *b, int nsteps) {
    if (nsteps < 0) return fail(LBM_ERR_ARG, "nsteps < 0");
    if (!b->collided && nsteps > 0) {
        int rc = eager_steps(b, 1);
        if (rc) return rc;
        --nsteps;
    }
    if (b->graph_steps > 0) {
        while (nsteps >= b->graph_steps) {
            if (!b->graph[b->cur]) {
                int rc = capture(b, b->cur);
                if (rc) return rc;
            }
            HIPCHK(hipGraphLaunch(b->graph[b->cur], b->stream));
            if (b->graph_flip[b->cur]) b->cur = 1 - b->cur;
            b->steps += b->graph_steps;
            nsteps -= b->graph_steps;
        }
    }
    return eager_steps(b, nsteps);
}

template <typename T> T *typed(void *p, int64_t off) { return static_cast<T *>(p) + off; }

}  // namespace

extern "C" {

int lbm_hip_batch_create(const lbm_problem *probs, int n, lbm_batch **out) {
    if (!out || !probs || n < 1) return fail(LBM_ERR_ARG, "bad batch request");
    *out = nullptr;
    int storage = probs[0].storage;
    if (storage < LBM_STORE_F64 || storage > LBM_STORE_F16)
        return fail(LBM_ERR_ARG, "unknown storage format");
    for (int c = 0; c < n; ++c) {
        if (probs[c].storage != storage)
            return fail(LBM_ERR_ARG, "mixed storage formats in one batch");
        const int rc = validate_problem(&probs[c]);
        if (rc) return fail(rc, "channel " + std::to_string(c) + ": " + lbm_g_err);
    }
    lbm_batch *b = new (std::nothrow) lbm_batch();
    if (!b) return fail(LBM_ERR_NOMEM, "host allocation");
    b->n = n;
    b->storage = storage;
    b->min_lx = probs[0].lx;
    for (int c = 1; c < n; ++c) b->min_lx = std::min(b->min_lx, probs[c].lx);
    b->depth = b->min_lx >= 4 ? 2 : 1;
    // fastest on the 512-channel workload: LDS for 3 / 4 / 5 workgroups per CU
    b->tile_cols = storage == LBM_STORE_F64 ? 8 : storage == LBM_STORE_F32 ? 14 : 10;
    b->ents.resize(n);
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        b->num_cu = prop.multiProcessorCount;
    // pool offsets (256-element aligned) and parameter slots
    int64_t off = 0, poff = 0;
    std::vector<int64_t> param_off(n);
    for (int c = 0; c < n; ++c) {
        const lbm_problem &p = probs[c];
        BatchEntry &e = b->ents[c];
        std::memset(&e, 0, sizeof e);
        e.p.lay = make_layout(p.lx, p.ly);
        e.p.ox = p.obst_x;
        e.p.oy = p.obst_y;
        e.p.r2 = p.obst_r >= 0 ? p.obst_r * p.obst_r : -1;
        e.p.omega = p.omega;
        e.p.zgb = p.zgb;
        e.off = off;
        off += (e.p.lay.span() + 255) / 256 * 256;
        param_off[c] = poff;
        poff += 4 * (int64_t)(p.lx + 2) + 4 * (int64_t)(p.ly + 2) + (p.ly + 2);
        b->cells += (int64_t)p.lx * p.ly;
        b->lx_of.push_back(p.lx);
        b->ly_of.push_back(p.ly);
    }
    b->pool_elems = off;
    auto bail = [&](hipError_t e, const char *what) {
        lbm_hip_batch_destroy(b);
        return hip_fail(e, what);
    };
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming);
    if (e != hipSuccess) return bail(e, "stream/event create");
    for (const void *k : {(const void *)k_batch_block<double>, (const void *)k_batch_block<float>,
                          (const void *)k_batch_block<_Float16>})
        if (e == hipSuccess)
            e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return bail(e, "hipFuncSetAttribute");
    for (int k = 0; k < 2; ++k) {
        e = hipMalloc(&b->pool[k], (size_t)off * esize(storage));
        if (e != hipSuccess) return bail(e, "hipMalloc pool");
    }
    e = hipMalloc((void **)&b->params, (size_t)poff * sizeof(double));
    if (e != hipSuccess) return bail(e, "hipMalloc params");
    e = hipMalloc((void **)&b->d_ents, n * sizeof(BatchEntry));
    if (e != hipSuccess) return bail(e, "hipMalloc entries");
    e = hipMalloc((void **)&b->d_step, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b->d_step, 0, sizeof(uint32_t));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return bail(e, "step counter");
    // per-channel side parameters
    std::vector<double> host(poff, 0.0);
    for (int c = 0; c < n; ++c) {
        const lbm_problem &p = probs[c];
        BatchEntry &ent = b->ents[c];
        int64_t q = param_off[c];
        for (int s = 0; s < 4; ++s) {
            int len = (s == LBM_SIDE_UPPER || s == LBM_SIDE_LOWER) ? p.lx + 2 : p.ly + 2;
            double *h0 = &host[q], *h1 = &host[q + len];
            if (p.side[s].p0) std::memcpy(h0, p.side[s].p0, len * sizeof(double));
            else if (p.side[s].kind == LBM_BC_PRESSURE) std::fill(h0, h0 + len, 1.0);
            if (p.side[s].p1) std::memcpy(h1, p.side[s].p1, len * sizeof(double));
            ent.p.side[s].kind = p.side[s].kind;
            ent.p.side[s].family = p.side[s].family;
            ent.p.side[s].p0 = b->params + q;
            ent.p.side[s].p1 = b->params + q + len;
            if (s == LBM_SIDE_RIGHT) {
                // outlet density record starts as the right side's p0
                std::memcpy(&host[param_off[c] + 4 * (int64_t)(p.lx + 2) + 4 * (int64_t)(p.ly + 2)],
                            h0, len * sizeof(double));
            }
            q += 2 * (int64_t)len;
        }
        ent.p.outlet_rho = b->params + q;
        ent.p.step = b->d_step;
    }
    e = hipMemcpy(b->params, host.data(), (size_t)poff * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(e, "params upload");
    int rc = build_items(b);
    if (rc) { std::string m = lbm_g_err; lbm_hip_batch_destroy(b); return fail(rc, m); }
    *out = b;
    return LBM_OK;
}

int lbm_hip_batch_destroy(lbm_batch *b) {
    if (!b) return LBM_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (int k = 0; k < 2; ++k) if (b->pool[k]) (void)hipFree(b->pool[k]);
    if (b->params) (void)hipFree(b->params);
    if (b->d_ents) (void)hipFree(b->d_ents);
    if (b->d_item_chan) (void)hipFree(b->d_item_chan);
    if (b->d_tile_chan) (void)hipFree(b->d_tile_chan);
    if (b->d_cw) (void)hipFree(b->d_cw);
    if (b->d_edge_items) (void)hipFree(b->d_edge_items);
    if (b->d_in_items) (void)hipFree(b->d_in_items);
    for (void *q : {(void *)b->d_dcw, (void *)b->d_din, (void *)b->d_frame[0],
                    (void *)b->d_frame[1], (void *)b->d_frame[2], (void *)b->d_sedge,
                    (void *)b->d_sin, b->mid[0], b->mid[1]})
        if (q) (void)hipFree(q);
    if (b->d_step) (void)hipFree(b->d_step);
    if (b->side) (void)hipStreamSynchronize(b->side);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return LBM_OK;
}

int lbm_hip_batch_init_equilibrium(lbm_batch *b, const double *const *init_ux) {
    if (!b || !init_ux) return fail(LBM_ERR_ARG, "null argument");
    HIPCHK(hipMemsetAsync(b->d_step, 0, sizeof(uint32_t), b->stream));
    // both pools to f = 0 (zero bits in FP64, -w in deviation storage)
    for (int k = 0; k < 2; ++k) {
        if (b->storage == LBM_STORE_F64) {
            HIPCHK(hipMemsetAsync(b->pool[k], 0, (size_t)b->pool_elems * 8, b->stream));
        } else {
            for (int c = 0; c < b->n; ++c)
                HIPCHK(launch_fill_zero_f(b->storage, b->ents[c].p.lay,
                                          (char *)b->pool[k] + b->ents[c].off * esize(b->storage),
                                          b->stream));
        }
    }
    double *d_ux = nullptr;
    int64_t tot = 0;
    for (int c = 0; c < b->n; ++c) tot += b->ly_of[c] + 2;
    HIPCHK(hipMalloc((void **)&d_ux, tot * sizeof(double)));
    std::vector<double> h(tot);
    int64_t q = 0;
    for (int c = 0; c < b->n; ++c) {
        std::memcpy(&h[q], init_ux[c], (b->ly_of[c] + 2) * sizeof(double));
        q += b->ly_of[c] + 2;
    }
    hipError_t e = hipMemcpy(d_ux, h.data(), tot * sizeof(double), hipMemcpyHostToDevice);
    q = 0;
    for (int c = 0; c < b->n && e == hipSuccess; ++c) {
        const Layout &L = b->ents[c].p.lay;
        void *dst = (char *)b->pool[0] + b->ents[c].off * esize(b->storage);
        e = b->storage == LBM_STORE_F64
                ? launch_init_f64(L, d_ux + q, (double *)dst, b->stream)
                : launch_init_lp(b->storage, L, d_ux + q, dst, b->stream);
        q += b->ly_of[c] + 2;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(d_ux);
    if (e != hipSuccess) return hip_fail(e, "batch init");
    b->cur = 0;
    b->collided = 0;
    b->steps = 0;
    return LBM_OK;
}

int lbm_hip_batch_set_launch(lbm_batch *b, int wgs_per_cu, int ty_cap, int graph_steps) {
    if (!b || wgs_per_cu < 0 || graph_steps < 0) return fail(LBM_ERR_ARG, "bad launch setting");
    HIPCHK(hipStreamSynchronize(b->stream));
    if (wgs_per_cu > 0) b->wgs_per_cu = wgs_per_cu;
    if (graph_steps != b->graph_steps) drop_graphs(b);
    b->graph_steps = graph_steps;
    if (ty_cap > 0) {
        if (ty_cap != 32 && ty_cap != 64 && ty_cap != 128 && ty_cap != 256)
            return fail(LBM_ERR_ARG, "ty_cap must be 32, 64, 128 or 256");
        if (ty_cap != b->ty_cap) {
            b->ty_cap = ty_cap;
            return build_items(b);
        }
    }
    drop_graphs(b);
    return LBM_OK;
}

int lbm_hip_batch_set_depth(lbm_batch *b, int depth, int tile_cols) {
    if (!b || (depth != 1 && depth != 2) || tile_cols < 0)
        return fail(LBM_ERR_ARG, "bad depth setting");
    if (tile_cols > 0 && (tile_cols < 3 || batch_block_lds(tile_cols, b->storage) > 160 * 1024))
        return fail(LBM_ERR_ARG, "tile_cols must be in [3, 33] (FP64) or [3, 69] (FP32, FP16)");
    // the outlet extrapolation of a tile's second step reads the first step
    // of column lx-3 (as the single-lattice two-step sweeps)
    if (depth == 2 && b->min_lx < 4)
        return fail(LBM_ERR_UNSUPPORTED, "two steps per launch need lx >= 4 in every channel");
    HIPCHK(hipStreamSynchronize(b->stream));
    drop_graphs(b);
    b->depth = depth;
    if (tile_cols > 0 && tile_cols != b->tile_cols) {
        b->tile_cols = tile_cols;
        return build_items(b);
    }
    return LBM_OK;
}

int lbm_hip_batch_set_sweep(lbm_batch *b, int variant, int chunk_cols) {
    if (!b || (variant != LBM_TIGHT_BLOCK && variant != LBM_PANEL && variant != LBM_PANEL4 &&
               variant != LBM_PANEL6))
        return fail(LBM_ERR_ARG, "a batch's multi-step sweep is LBM_TIGHT_BLOCK, LBM_PANEL, "
                                 "LBM_PANEL4 or LBM_PANEL6");
    if (chunk_cols != 0 && (chunk_cols < 4 || chunk_cols > 64 || chunk_cols % 2))
        return fail(LBM_ERR_ARG, "chunk_cols must be even, in [4, 64] (0: keep)");
    const bool deep = variant == LBM_PANEL4 || variant == LBM_PANEL6;
    const int levels = sweep_levels(variant);
    const int cx = chunk_cols ? chunk_cols : deep ? b->deep_cx : b->panel_cx;
    if (variant == b->sweep && cx == (deep ? b->deep_cx : b->panel_cx)) return LBM_OK;
    HIPCHK(hipStreamSynchronize(b->stream));
    drop_graphs(b);
    if (variant != LBM_TIGHT_BLOCK && (!b->d_cw || (!deep && cx != b->panel_cx))) {
        const int old = b->panel_cx;
        if (!deep) b->panel_cx = cx;
        const int rc = build_waves(b);
        if (rc) {
            b->panel_cx = old;
            if (rc != LBM_ERR_UNSUPPORTED && rc != LBM_ERR_ARG) {
                // a device allocation failed part-way: the two-step tables are gone
                if (b->d_cw) (void)hipFree(b->d_cw);
                b->d_cw = nullptr;
                b->nedge_items = b->nin_wgs = 0;
                b->deep_levels = 0;
                b->sweep = LBM_TIGHT_BLOCK;
            }
            return rc;
        }
        b->deep_levels = 0;   // deep tables, if any, list the old two-step grids
    }
    if (deep && (!b->d_dcw || cx != b->deep_cx || levels != b->deep_levels)) {
        b->deep_cx = cx;
        b->deep_levels = levels;
        const int rc = build_deep(b);
        if (rc) {
            // the deep tables are gone: fall back to the two-step march
            b->deep_levels = 0;
            b->sweep = LBM_PANEL;
            return rc;
        }
    }
    if (!deep) b->panel_cx = cx;
    b->sweep = variant;
    return LBM_OK;
}

int lbm_hip_batch_step(lbm_batch *b, int nsteps) {
    if (!b) return fail(LBM_ERR_ARG, "null batch");
    return step_impl(b, nsteps);
}

int lbm_hip_batch_sync(lbm_batch *b) {
    if (!b) return fail(LBM_ERR_ARG, "null batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_hip_batch_time(lbm_batch *b, int nsteps, double *ms) {
    if (!b || !ms) return fail(LBM_ERR_ARG, "null argument");
    if (!b->collided) {
        int rc = step_impl(b, 1);
        if (rc) return rc;
    }
    hipEvent_t a, z;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&z));
    hipError_t e = hipEventRecord(a, b->stream);
    int rc = e == hipSuccess ? step_impl(b, nsteps) : LBM_OK;
    if (!rc && e == hipSuccess) e = hipEventRecord(z, b->stream);
    if (!rc && e == hipSuccess) e = hipEventSynchronize(z);
    float t = 0.f;
    if (!rc && e == hipSuccess) e = hipEventElapsedTime(&t, a, z);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(z);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "batch timing");
    *ms = t;
    return LBM_OK;
}

int lbm_hip_batch_download(lbm_batch *b, int c, double *lattice, double *tmp,
                           double *outlet_rho) {
    if (!b || c < 0 || c >= b->n) return fail(LBM_ERR_ARG, "bad channel index");
    const BatchEntry &e = b->ents[c];
    const Layout &L = e.p.lay;
    size_t es = esize(b->storage);
    void *P = (char *)b->pool[b->cur] + e.off * es;
    void *Q = (char *)b->pool[1 - b->cur] + e.off * es;
    void *lat_src = P, *tmp_src = Q;
    if (b->collided) {
        HIPCHK(b->storage == LBM_STORE_F64
                   ? launch_stream_f64(L, (const double *)P, (double *)Q, b->stream)
                   : launch_stream_lp(b->storage, L, P, Q, b->stream));
        lat_src = Q;
        tmp_src = P;
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    double *wide = nullptr;
    if (b->storage != LBM_STORE_F64)
        HIPCHK(hipMalloc((void **)&wide, (size_t)L.span() * sizeof(double)));
    int rc = LBM_OK;
    void *srcs[2] = {lat_src, tmp_src};
    double *hosts[2] = {lattice, tmp};
    for (int k = 0; k < 2 && !rc; ++k) {
        if (!hosts[k]) continue;
        double *dev = (double *)srcs[k];
        if (wide) {
            hipError_t ee = launch_widen(b->storage, L, srcs[k], wide, b->stream);
            if (ee == hipSuccess) ee = hipStreamSynchronize(b->stream);
            if (ee != hipSuccess) { rc = hip_fail(ee, "widen"); break; }
            dev = wide;
        }
        rc = transfer_f64(L, hosts[k], dev, false);
    }
    if (wide) (void)hipFree(wide);
    if (!rc && outlet_rho) {
        hipError_t ee = hipMemcpy(outlet_rho, e.p.outlet_rho, (L.ly + 2) * sizeof(double),
                                  hipMemcpyDeviceToHost);
        if (ee != hipSuccess) rc = hip_fail(ee, "outlet download");
    }
    return rc;
}

int lbm_hip_batch_macros(lbm_batch *b, int c, double *rho, double *ux, double *uy) {
    if (!b || c < 0 || c >= b->n) return fail(LBM_ERR_ARG, "bad channel index");
    const BatchEntry &e = b->ents[c];
    const Layout &L = e.p.lay;
    size_t n = (size_t)L.lx * L.ly;
    double *d = nullptr;
    HIPCHK(hipMalloc((void **)&d, 3 * n * sizeof(double)));
    void *src = (char *)b->pool[b->cur] + e.off * esize(b->storage);
    hipError_t er = b->storage == LBM_STORE_F64
                        ? launch_macros_f64(L, 1, L.lx, b->collided, (const double *)src, d,
                                            d + n, d + 2 * n, b->stream)
                        : launch_macros_lp(b->storage, L, 1, L.lx, b->collided, src, d, d + n,
                                           d + 2 * n, b->stream);
    if (er == hipSuccess) er = hipStreamSynchronize(b->stream);
    if (er == hipSuccess && rho) er = hipMemcpy(rho, d, n * sizeof(double), hipMemcpyDeviceToHost);
    if (er == hipSuccess && ux) er = hipMemcpy(ux, d + n, n * sizeof(double), hipMemcpyDeviceToHost);
    if (er == hipSuccess && uy)
        er = hipMemcpy(uy, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (er != hipSuccess) return hip_fail(er, "batch macros");
    return LBM_OK;
}

int64_t lbm_hip_batch_cells(const lbm_batch *b) { return b ? b->cells : -1; }
int64_t lbm_hip_batch_steps_done(const lbm_batch *b) { return b ? b->steps : -1; }

size_t lbm_hip_batch_footprint(const lbm_batch *b) {
    if (!b) return 0;
    const int bufs = 2 + (b->mid[0] != nullptr) + (b->mid[1] != nullptr);
    return bufs * (size_t)b->pool_elems * esize(b->storage) +
           (size_t)(b->nitems + b->ntiles) * sizeof(int) + b->ents.size() * sizeof(BatchEntry);
}

}  // extern "C"
