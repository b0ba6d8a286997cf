// capi.hip -- the extern "C" surface of include/lbm_hip.h: device contexts,
// state bookkeeping between the two state conventions, host transfers and
// the variant dispatch.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "launch.hpp"
#include "host_util.hpp"

using namespace lbm;

thread_local std::string lbm_g_err;

namespace {
size_t elem_size(int storage) { return storage == LBM_STORE_F64 ? 8 : storage == LBM_STORE_F32 ? 4 : 2; }
constexpr int MAX_DEVICES = 64;
// Stream-ordered memory pools of the library, per device (lbm_hip_pool_reserve)
std::mutex g_pool_mu;
hipMemPool_t g_pool[MAX_DEVICES];
int g_pool_users[MAX_DEVICES];
// the moments scratch of a context holds at most this many nodes' fields
constexpr int64_t FIELDS_CHUNK_NODES = 4 << 20;
int64_t fields_nodes_of(int lx, int ly) {
    const int64_t n = (int64_t)lx * ly;
    return n < FIELDS_CHUNK_NODES ? n : (FIELDS_CHUNK_NODES / ly > 0 ? FIELDS_CHUNK_NODES / ly : 1) * (int64_t)ly;
}
}  // namespace

// Staging of lbm_hip_apply_node_op_noalloc: LBM_NODE_OP_MAX population
// vectors of 9 doubles, then their p0 and their p1 values.  Each GPU holds
// its own copy at its own address; host code reaches it only through the
// address hipGetSymbolAddress returns on that device (the array's name in
// host code is a stand-in, not a device address).
static __device__ __align__(256) double scratch[LBM_NODE_OP_MAX * 11];

namespace {
std::mutex g_scratch_mu[MAX_DEVICES];      // one per device: serialises the calls
double *g_scratch_ptr[MAX_DEVICES];        // device address of `scratch`, per device
}  // namespace

struct lbm_ctx {
    Problem prob;
    int storage;
    void *buf[2];
    int cur;         // buffer holding L (streamed) or P (collided)
    int collided;    // state convention of buf[cur]
    int64_t steps;
    double *d_p0[4], *d_p1[4];
    double *d_outlet;
    double *d_partial;
    double *d_out4;
    int64_t npartial;
    hipStream_t stream;
    int tile_x, tile_y;
    size_t footprint;
    int graph_steps;               // fused steps per captured graph (0: eager)
    uint32_t *d_step;              // collision counter (keys FP32 stochastic rounding)
    void *buf3, *buf4, *buf5;      // further population buffers of the four-, six- and eight-step launches
    hipStream_t side;              // second stream for the panel's edge waves
    hipEvent_t fork, join;
    hipGraphExec_t graph[2];       // fused steps, by starting buffer
    hipGraphExec_t pgraph[2];      // panel launches, by starting buffer
    // launch timing of the kernels that carry the lattice updates
    int timing;                    // record them (lbm_hip_kernel_timing)
    int capturing;                 // inside a graph capture: nothing is recorded
    std::vector<hipEvent_t> tev;   // event pool, two per recorded launch
    struct Timed { int depth; std::string name; double nodes; };
    std::vector<Timed> tlog;
    KernelProbe probe;
    hipMemPool_t pool;             // the device's library pool, or null (hipMalloc)
    double *d_init;                // ly + 2 initial / inlet velocities
    double *d_fields;              // moments scratch: 3 x fields_nodes doubles
    int64_t fields_nodes;
};

namespace {

int side_len(const lbm_problem *p, int s) {
    return (s == LBM_SIDE_UPPER || s == LBM_SIDE_LOWER) ? p->lx + 2 : p->ly + 2;
}

int upload_side(lbm_ctx *c, int s, const lbm_problem *p, const double *p0, const double *p1) {
    int n = side_len(p, s);
    std::vector<double> a(n, 0.0), b(n, 0.0);
    if (p0) std::memcpy(a.data(), p0, n * sizeof(double));
    else if (p->side[s].kind == LBM_BC_PRESSURE) std::fill(a.begin(), a.end(), 1.0);
    if (p1) std::memcpy(b.data(), p1, n * sizeof(double));
    // copies on the context's stream, waited for before the host arrays go
    HIPCHK(hipMemcpyAsync(c->d_p0[s], a.data(), n * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipMemcpyAsync(c->d_p1[s], b.data(), n * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
    if (s == LBM_SIDE_RIGHT)
        HIPCHK(hipMemcpyAsync(c->d_outlet, a.data(), n * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

double *dbuf(lbm_ctx *c, int k) { return static_cast<double *>(c->buf[k]); }

// Device allocations of a context: stream-ordered on the context's stream
// from the device's library pool when lbm_hip_pool_reserve made one (no
// device-wide synchronisation when contexts come and go), else hipMalloc /
// hipFree.
hipError_t ctx_malloc(lbm_ctx *c, void **p, size_t bytes) {
    if (c->pool) return hipMallocFromPoolAsync(p, bytes, c->pool, c->stream);
    return hipMalloc(p, bytes);
}

void ctx_free(lbm_ctx *c, void *p) {
    if (!p) return;
    if (c->pool) (void)hipFreeAsync(p, c->stream);
    else (void)hipFree(p);
}

const char *storage_name(int storage) {
    return storage == LBM_STORE_F64 ? "double" : storage == LBM_STORE_F32 ? "float" : "_Float16";
}

// Launch timing: a probe for the next launch of a dominant kernel, or null
// when timing is off, a graph is being captured or the record is full.
constexpr size_t MAX_TIMED = 4096;

KernelProbe *probe_for(lbm_ctx *c) {
    if (!c->timing || c->capturing || c->tlog.size() >= MAX_TIMED) return nullptr;
    const size_t k = 2 * c->tlog.size();
    while (c->tev.size() < k + 2) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->tev.push_back(e);
    }
    c->probe = KernelProbe{};
    c->probe.t0 = c->tev[k];
    c->probe.t1 = c->tev[k + 1];
    return &c->probe;
}

// keep the probe's launch if the launcher recorded one (a sweep without
// interior waves launches no interior kernel)
void probe_done(lbm_ctx *c, KernelProbe *pr, int depth) {
    if (pr && pr->name[0]) c->tlog.push_back({depth, pr->name, pr->nodes});
}

// storage dispatch of the one-step kernels
hipError_t run_collide(lbm_ctx *c, bool pull, int from, int to) {
    KernelProbe *pr = probe_for(c);
    if (pr) {
        std::snprintf(pr->name, sizeof pr->name, "k_collide<%s, %s, %d>",
                      storage_name(c->storage), pull ? "true" : "false",
                      one_step_rows(c->prob.lay, c->storage));
        pr->nodes = (double)c->prob.lay.lx * c->prob.lay.ly;
        hipError_t e0 = hipEventRecord(pr->t0, c->stream);
        if (e0 != hipSuccess) return e0;
    }
    hipError_t e = c->storage == LBM_STORE_F64
                       ? launch_collide_f64(c->prob, pull, dbuf(c, from), dbuf(c, to), c->stream)
                       : launch_collide_lp(c->storage, c->prob, pull, c->buf[from], c->buf[to],
                                           c->stream);
    if (e == hipSuccess && pr) {
        e = hipEventRecord(pr->t1, c->stream);
        probe_done(c, pr, 1);
    }
    if (e == hipSuccess && c->storage != LBM_STORE_F64) e = launch_bump(c->d_step, 1, c->stream);
    return e;
}

hipError_t run_stream(lbm_ctx *c, int from, int to) {
    if (c->storage == LBM_STORE_F64)
        return launch_stream_f64(c->prob.lay, dbuf(c, from), dbuf(c, to), c->stream);
    return launch_stream_lp(c->storage, c->prob.lay, c->buf[from], c->buf[to], c->stream);
}

// Bring the context into the collided convention (P = buf[cur]).
int to_collided(lbm_ctx *c) {
    if (c->collided) return LBM_OK;
    // one collision of step m from the streamed state: L -> T
    HIPCHK(run_collide(c, false, c->cur, 1 - c->cur));
    c->cur = 1 - c->cur;
    c->collided = 1;
    c->steps += 1;  // the collided state after this represents step m+1
    return LBM_OK;
}

// Bring the context into the streamed convention (L = buf[cur]).
int to_streamed(lbm_ctx *c) {
    if (!c->collided) return LBM_OK;
    HIPCHK(run_stream(c, c->cur, 1 - c->cur));
    c->cur = 1 - c->cur;
    c->collided = 0;
    return LBM_OK;
}

void drop_graphs(lbm_ctx *c) {
    for (int k = 0; k < 2; ++k) {
        if (c->graph[k]) {
            (void)hipGraphExecDestroy(c->graph[k]);
            c->graph[k] = nullptr;
        }
        if (c->pgraph[k]) {
            (void)hipGraphExecDestroy(c->pgraph[k]);
            c->pgraph[k] = nullptr;
        }
    }
}

// one two-step panel launch from buffer `from` into 1 - from
hipError_t panel_launch(lbm_ctx *c, int from) {
    WaveGrid g = wave_grid(c->prob.lay, c->tile_x);
    KernelProbe *pr = probe_for(c);
    hipError_t e = launch_wpanel(c->storage, c->prob, g, c->buf[from], c->buf[1 - from], c->stream,
                                 c->side, c->fork, c->join, pr);
    if (e == hipSuccess) probe_done(c, pr, 2);
    if (e == hipSuccess && c->storage != LBM_STORE_F64) e = launch_bump(c->d_step, 2, c->stream);
    return e;
}

// one launch of `levels` (4, 6 or 8) steps from buffer `from` into 1 - from,
// the perimeter frame passing through buf3 (and buf4, buf5)
hipError_t deep_launch(lbm_ctx *c, int from, int levels) {
    void *mid[3] = {c->buf3, c->buf4, c->buf5};
    KernelProbe *pr = probe_for(c);
    hipError_t e = launch_wpanel_deep(c->storage, c->prob, c->tile_x, levels, c->buf[from], mid,
                                      c->buf[1 - from], c->stream, c->side, c->fork, c->join, pr);
    if (e == hipSuccess) probe_done(c, pr, levels);
    if (e == hipSuccess && c->storage != LBM_STORE_F64)
        e = launch_bump(c->d_step, (uint32_t)levels, c->stream);
    return e;
}

// allocate a further population buffer on first use (counted in the footprint)
int extra_buffer(lbm_ctx *c, void **b) {
    if (*b) return LBM_OK;
    size_t bytes = (size_t)c->prob.lay.span() * elem_size(c->storage);
    HIPCHK(ctx_malloc(c, b, bytes));
    c->footprint += bytes;
    return LBM_OK;
}

// capture graph_steps / 2 panel launches starting from buffer `start` (the
// perimeter stream joins the capture through the fork / join events)
int capture_panel(lbm_ctx *c, int start) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    hipError_t e = hipSuccess;
    int cur = start;
    c->capturing = 1;
    for (int k = 0; k < c->graph_steps / 2 && e == hipSuccess; ++k) {
        e = panel_launch(c, cur);
        cur = 1 - cur;
    }
    c->capturing = 0;
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return hip_fail(e, "capture"); }
    if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
    hipError_t e3 = hipGraphInstantiate(&c->pgraph[start], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e3 != hipSuccess) return hip_fail(e3, "hipGraphInstantiate");
    return LBM_OK;
}

// capture graph_steps fused pull steps starting from buffer `start`
int capture_fused(lbm_ctx *c, int start) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    hipError_t e = hipSuccess;
    int cur = start;
    c->capturing = 1;
    for (int k = 0; k < c->graph_steps && e == hipSuccess; ++k) {
        e = run_collide(c, true, cur, 1 - cur);
        cur = 1 - cur;
    }
    c->capturing = 0;
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return hip_fail(e, "capture"); }
    if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
    hipError_t e3 = hipGraphInstantiate(&c->graph[start], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e3 != hipSuccess) return hip_fail(e3, "hipGraphInstantiate");
    return LBM_OK;
}

int step_impl(lbm_ctx *c, int variant, int nsteps) {
    if (nsteps < 0) return fail(LBM_ERR_ARG, "nsteps < 0");
    if (nsteps == 0) return LBM_OK;
    switch (variant) {
    case LBM_ORIGIN: {
        // collide (lb.c:172-201) then stream (lb.c:246-287), every step
        int rc = to_streamed(c);
        if (rc) return rc;
        for (int k = 0; k < nsteps; ++k) {
            HIPCHK(run_collide(c, false, c->cur, 1 - c->cur));
            HIPCHK(run_stream(c, 1 - c->cur, c->cur));
            c->steps += 1;
        }
        return LBM_OK;
    }
    case LBM_COMBINE:
    case LBM_TIGHT: {
        int k = 0;
        if (!c->collided) {
            int rc = to_collided(c);
            if (rc) return rc;
            k = 1;
        }
        if (c->graph_steps > 0) {
            while (nsteps - k >= c->graph_steps) {
                if (!c->graph[c->cur]) {
                    int rc = capture_fused(c, c->cur);
                    if (rc) return rc;
                }
                HIPCHK(hipGraphLaunch(c->graph[c->cur], c->stream));
                if (c->graph_steps & 1) c->cur = 1 - c->cur;
                c->steps += c->graph_steps;
                k += c->graph_steps;
            }
        }
        for (; k < nsteps; ++k) {
            HIPCHK(run_collide(c, true, c->cur, 1 - c->cur));
            c->cur = 1 - c->cur;
            c->steps += 1;
        }
        return LBM_OK;
    }
    case LBM_PANEL4:
    case LBM_PANEL6: {
        // four (six) steps per launch while four (six) remain; the rest with
        // the shallower sweeps
        if (c->prob.lay.lx < 4)
            return fail(LBM_ERR_UNSUPPORTED, "two-step variants need lx >= 4");
        const int levels = variant == LBM_PANEL6 ? 6 : 4;
        const int fallback = variant == LBM_PANEL6 ? LBM_PANEL4 : LBM_PANEL;
        if (!wpanel_deep_possible(c->prob.lay, c->tile_x, c->storage, levels))
            return step_impl(c, fallback, nsteps);
        int rc = extra_buffer(c, &c->buf3);
        if (!rc && levels == 6) rc = extra_buffer(c, &c->buf4);
        if (rc) return rc;
        int k = 0;
        if (!c->collided) {
            rc = to_collided(c);
            if (rc) return rc;
            k = 1;
        }
        // launches of `levels` steps while they fit; a six-step sweep whose
        // remainder would need a two-step launch (2 or 3 steps left) runs one
        // six-step launch fewer and one eight-step launch instead (8 = 8, 9 =
        // 8 + 1), or two four-step launches where the lattice has no interior
        // for eight steps (8 = 4 + 4): on 16384 x 4096 a two-step launch costs
        // 1.0 ms per step, a four-step one 0.57, a six-step one 0.45
        int deep = (nsteps - k) / levels;
        const int rest = (nsteps - k) % levels;
        int eight = 0;
        if (levels == 6 && deep > 0 && (rest == 2 || rest == 3)) {
            --deep;
            if (wpanel_deep_possible(c->prob.lay, c->tile_x, c->storage, 8)) {
                rc = extra_buffer(c, &c->buf5);
                if (rc) return rc;
                eight = 1;
            }
        }
        for (int j = 0; j < deep; ++j, k += levels) {
            HIPCHK(deep_launch(c, c->cur, levels));
            c->cur = 1 - c->cur;
            c->steps += levels;
        }
        if (eight) {
            HIPCHK(deep_launch(c, c->cur, 8));
            c->cur = 1 - c->cur;
            c->steps += 8;
            k += 8;
        }
        return k < nsteps ? step_impl(c, fallback, nsteps - k) : LBM_OK;
    }
    case LBM_TIGHT_BLOCK:
    case LBM_PANEL: {
        // two steps per launch (lb.c:936-1069, 1402-1538); an odd count ends
        // with one fused single step
        if (c->prob.lay.lx < 4)
            return fail(LBM_ERR_UNSUPPORTED, "two-step variants need lx >= 4");
        if (variant == LBM_PANEL && !wave_grid_fits(c->prob.lay))
            return fail(LBM_ERR_UNSUPPORTED, "lattice too tall for the panel sweeps");
        int k = 0;
        if (!c->collided) {
            int rc = to_collided(c);
            if (rc) return rc;
            k = 1;
        }
        const int gsteps = 2 * (c->graph_steps / 2);
        if (variant == LBM_PANEL && gsteps > 0) {
            while (nsteps - k >= gsteps) {
                if (!c->pgraph[c->cur]) {
                    int rc = capture_panel(c, c->cur);
                    if (rc) return rc;
                }
                HIPCHK(hipGraphLaunch(c->pgraph[c->cur], c->stream));
                if ((gsteps / 2) & 1) c->cur = 1 - c->cur;
                c->steps += gsteps;
                k += gsteps;
            }
        }
        for (; k + 2 <= nsteps; k += 2) {
            if (variant == LBM_PANEL) {
                HIPCHK(panel_launch(c, c->cur));
            } else {
                BlockGrid g = block_grid(c->prob.lay, c->tile_x,
                                         c->tile_y > 0 ? c->tile_y + 2 : 0,
                                         c->storage == LBM_STORE_F64 ? 8 : 4);
                KernelProbe *pr = probe_for(c);
                if (pr) {
                    std::snprintf(pr->name, sizeof pr->name, "k_block<%s, %d>",
                                  storage_name(c->storage), g.typ);
                    pr->nodes = (double)c->prob.lay.lx * c->prob.lay.ly;
                    HIPCHK(hipEventRecord(pr->t0, c->stream));
                }
                HIPCHK(launch_block(c->storage, c->prob, g, c->buf[c->cur], c->buf[1 - c->cur],
                                    c->stream));
                if (pr) {
                    HIPCHK(hipEventRecord(pr->t1, c->stream));
                    probe_done(c, pr, 2);
                }
                if (c->storage != LBM_STORE_F64) HIPCHK(launch_bump(c->d_step, 2, c->stream));
            }
            c->cur = 1 - c->cur;
            c->steps += 2;
        }
        if (k < nsteps) {
            HIPCHK(run_collide(c, true, c->cur, 1 - c->cur));
            c->cur = 1 - c->cur;
            c->steps += 1;
        }
        return LBM_OK;
    }
    default:
        return fail(LBM_ERR_ARG, "unknown variant");
    }
}

// host node layout (FP64) <-> device buffer k of any storage
int transfer(lbm_ctx *c, double *host, int k, bool to_device) {
    if (c->storage == LBM_STORE_F64) return transfer_f64(c->prob.lay, host, dbuf(c, k), to_device);
    const Layout &L = c->prob.lay;
    double *tmp = nullptr;
    HIPCHK(hipMalloc((void **)&tmp, (size_t)L.span() * sizeof(double)));
    hipError_t e = launch_widen(c->storage, L, c->buf[k], tmp, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    int rc = e == hipSuccess ? transfer_f64(c->prob.lay, host, tmp, to_device) : hip_fail(e, "widen");
    if (!rc && to_device) {
        e = launch_narrow(c->storage, L, tmp, c->buf[k], c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "narrow");
    }
    (void)hipFree(tmp);
    return rc;
}

}  // namespace

extern "C" {

const char *lbm_hip_last_error(void) { return lbm_g_err.c_str(); }

size_t lbm_hip_footprint_of(const lbm_problem *p) {
    if (!p || p->lx < 1 || p->ly < 1) return 0;
    Layout L = make_layout(p->lx, p->ly);
    size_t pops = 2 * (size_t)L.span() * elem_size(p->storage);
    size_t sides = 2 * sizeof(double) * (2 * (size_t)(p->lx + 2) + 2 * (size_t)(p->ly + 2));
    size_t diag = 4 * sizeof(double) * (size_t)diag_partials_needed(L) + 64;
    size_t fields = 3 * sizeof(double) * (size_t)fields_nodes_of(p->lx, p->ly);
    return pops + sides + 2 * (size_t)(p->ly + 2) * sizeof(double) + diag + fields;
}

size_t lbm_hip_footprint_for(const lbm_problem *p, int variant) {
    const size_t base = lbm_hip_footprint_of(p);
    if (!base) return 0;
    // the six-step sweep's eight-step launch (for remainders of 2 or 3
    // steps) passes its frame through a third further buffer
    const size_t extra = variant == LBM_PANEL6 ? 3 : variant == LBM_PANEL4 ? 1 : 0;
    return base + extra * (size_t)make_layout(p->lx, p->ly).span() * elem_size(p->storage);
}

int lbm_hip_create(const lbm_problem *p, lbm_ctx **out) {
    if (!out) return fail(LBM_ERR_ARG, "null output pointer");
    *out = nullptr;
    int rc = validate_problem(p);
    if (rc) return rc;
    lbm_ctx *c = new (std::nothrow) lbm_ctx();
    if (!c) return fail(LBM_ERR_NOMEM, "host allocation");
    c->storage = p->storage;
    c->prob.lay = make_layout(p->lx, p->ly);
    c->prob.ox = p->obst_x;
    c->prob.oy = p->obst_y;
    c->prob.r2 = p->obst_r >= 0 ? p->obst_r * p->obst_r : -1;
    c->prob.omega = p->omega;
    c->prob.zgb = p->zgb;
    c->npartial = diag_partials_needed(c->prob.lay);
    static bool lds_attr_set = false;
    if (!lds_attr_set) {
        hipError_t ea = block_kernel_setup();
        if (ea != hipSuccess) { delete c; return hip_fail(ea, "hipFuncSetAttribute"); }
        lds_attr_set = true;
    }
    size_t bytes = (size_t)c->prob.lay.span() * elem_size(p->storage);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return hip_fail(e, "hipStreamCreate"); }
    {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < MAX_DEVICES) {
            std::lock_guard<std::mutex> lock(g_pool_mu);
            if (g_pool[dev]) {
                c->pool = g_pool[dev];
                ++g_pool_users[dev];
            }
        }
    }
    e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
    if (e != hipSuccess) { lbm_hip_destroy(c); return hip_fail(e, "stream/event create"); }
    auto cleanup = [&](hipError_t err, const char *what) {
        lbm_hip_destroy(c);
        return hip_fail(err, what);
    };
    // everything below is ordered on the context's own stream: nothing here
    // waits for other contexts' work
    for (int k = 0; k < 2; ++k) {
        e = ctx_malloc(c, &c->buf[k], bytes);
        if (e != hipSuccess) return cleanup(e, "allocating populations");
        e = p->storage == LBM_STORE_F64
                ? hipMemsetAsync(c->buf[k], 0, bytes, c->stream)
                : launch_fill_zero_f(p->storage, c->prob.lay, c->buf[k], c->stream);
        if (e != hipSuccess) return cleanup(e, "clearing populations");
    }
    for (int s = 0; s < 4; ++s) {
        int n = side_len(p, s);
        e = ctx_malloc(c, (void **)&c->d_p0[s], n * sizeof(double));
        if (e != hipSuccess) return cleanup(e, "allocating side parameters");
        e = ctx_malloc(c, (void **)&c->d_p1[s], n * sizeof(double));
        if (e != hipSuccess) return cleanup(e, "allocating side parameters");
    }
    e = ctx_malloc(c, (void **)&c->d_outlet, (p->ly + 2) * sizeof(double));
    if (e == hipSuccess) e = ctx_malloc(c, (void **)&c->d_init, (p->ly + 2) * sizeof(double));
    if (e != hipSuccess) return cleanup(e, "allocating outlet");
    e = ctx_malloc(c, (void **)&c->d_partial, 4 * sizeof(double) * c->npartial);
    if (e != hipSuccess) return cleanup(e, "allocating partials");
    e = ctx_malloc(c, (void **)&c->d_out4, 4 * sizeof(double));
    if (e != hipSuccess) return cleanup(e, "allocating diagnostics");
    e = ctx_malloc(c, (void **)&c->d_step, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemsetAsync(c->d_step, 0, sizeof(uint32_t), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return cleanup(e, "step counter");
    for (int s = 0; s < 4; ++s) {
        rc = upload_side(c, s, p, p->side[s].p0, p->side[s].p1);
        if (rc) { std::string m = lbm_g_err; lbm_hip_destroy(c); return fail(rc, m); }
        c->prob.side[s].kind = p->side[s].kind;
        c->prob.side[s].family = p->side[s].family;
        c->prob.side[s].p0 = c->d_p0[s];
        c->prob.side[s].p1 = c->d_p1[s];
    }
    c->prob.outlet_rho = c->d_outlet;
    c->prob.step = c->d_step;
    c->cur = 0;
    c->collided = 0;
    c->steps = 0;
    c->footprint = lbm_hip_footprint_of(p);
    *out = c;
    return LBM_OK;
}

int lbm_hip_destroy(lbm_ctx *c) {
    if (!c) return LBM_OK;
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (int k = 0; k < 2; ++k) ctx_free(c, c->buf[k]);
    for (int s = 0; s < 4; ++s) {
        ctx_free(c, c->d_p0[s]);
        ctx_free(c, c->d_p1[s]);
    }
    for (void *p : {(void *)c->d_outlet, (void *)c->d_init, (void *)c->d_partial,
                    (void *)c->d_out4, (void *)c->d_step, c->buf3, c->buf4, c->buf5,
                    (void *)c->d_fields})
        ctx_free(c, p);
    if (c->pool) {
        // the pool's frees are ordered on the stream: let them land first
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < MAX_DEVICES) {
            std::lock_guard<std::mutex> lock(g_pool_mu);
            if (g_pool[dev] == c->pool && g_pool_users[dev] > 0) --g_pool_users[dev];
        }
    }
    for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LBM_OK;
}

int lbm_hip_init_equilibrium(lbm_ctx *c, const double *init_ux) {
    if (!c || !init_ux) return fail(LBM_ERR_ARG, "null argument");
    const Layout &L = c->prob.lay;
    size_t bytes = (size_t)L.span() * elem_size(c->storage);
    double *d_ux = c->d_init;
    hipError_t e = hipMemcpyAsync(d_ux, init_ux, (L.ly + 2) * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream);
    // both buffers start as f = 0 (calloc, lb.c:89-92); in deviation storage
    // that is g = -w, not a zero bit pattern
    for (int k = 0; k < 2 && e == hipSuccess; ++k)
        e = c->storage == LBM_STORE_F64 ? hipMemsetAsync(c->buf[k], 0, bytes, c->stream)
                                        : launch_fill_zero_f(c->storage, L, c->buf[k],
                                                             c->stream);
    if (e == hipSuccess)
        e = c->storage == LBM_STORE_F64 ? launch_init_f64(L, d_ux, dbuf(c, 0), c->stream)
                                        : launch_init_lp(c->storage, L, d_ux, c->buf[0], c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->d_step, 0, sizeof(uint32_t), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "init_equilibrium");
    c->cur = 0;
    c->collided = 0;
    c->steps = 0;
    return LBM_OK;
}

int lbm_hip_upload(lbm_ctx *c, const double *lattice, const double *tmp,
                   const double *outlet_rho) {
    if (!c || !lattice || !tmp) return fail(LBM_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    int rc = transfer(c, const_cast<double *>(lattice), 0, true);
    if (!rc) rc = transfer(c, const_cast<double *>(tmp), 1, true);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(c->d_step, 0, sizeof(uint32_t), c->stream));
    if (outlet_rho) {
        HIPCHK(hipMemcpy(c->d_outlet, outlet_rho, (c->prob.lay.ly + 2) * sizeof(double),
                         hipMemcpyHostToDevice));
        if (!c->prob.zgb)
            HIPCHK(hipMemcpy(c->d_p0[LBM_SIDE_RIGHT], outlet_rho,
                             (c->prob.lay.ly + 2) * sizeof(double), hipMemcpyHostToDevice));
    }
    c->cur = 0;
    c->collided = 0;
    c->steps = 0;
    return LBM_OK;
}

int lbm_hip_download(lbm_ctx *c, double *lattice, double *tmp, double *outlet_rho) {
    if (!c) return fail(LBM_ERR_ARG, "null context");
    int lat_buf = c->cur, tmp_buf = 1 - c->cur;
    if (c->collided) {
        // post-stream state into the free buffer; P itself is left as it is
        HIPCHK(run_stream(c, c->cur, 1 - c->cur));
        lat_buf = 1 - c->cur;
        tmp_buf = c->cur;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    int rc = LBM_OK;
    if (lattice) rc = transfer(c, lattice, lat_buf, false);
    if (!rc && tmp) rc = transfer(c, tmp, tmp_buf, false);
    if (rc) return rc;
    if (outlet_rho)
        HIPCHK(hipMemcpy(outlet_rho, c->d_outlet, (c->prob.lay.ly + 2) * sizeof(double),
                         hipMemcpyDeviceToHost));
    return LBM_OK;
}

int lbm_hip_step(lbm_ctx *c, int variant, int nsteps) {
    if (!c) return fail(LBM_ERR_ARG, "null context");
    return step_impl(c, variant, nsteps);
}

int lbm_hip_collide(lbm_ctx *c) {
    if (!c) return fail(LBM_ERR_ARG, "null context");
    if (c->collided) return fail(LBM_ERR_STATE, "collide needs the streamed state (propagate first)");
    return to_collided(c);
}

int lbm_hip_propagate(lbm_ctx *c) {
    if (!c) return fail(LBM_ERR_ARG, "null context");
    if (!c->collided) return fail(LBM_ERR_STATE, "propagate needs a collided state (collide first)");
    return to_streamed(c);
}

int lbm_hip_make_periodic(lbm_ctx *c) {
    if (!c) return fail(LBM_ERR_ARG, "null context");
    // the copy acts on the post-stream state: L = buf[cur], T = buf[1 - cur]
    int rc = to_streamed(c);
    if (rc) return rc;
    HIPCHK(launch_periodic(c->storage, c->prob.lay, c->buf[1 - c->cur], c->buf[c->cur],
                           c->stream));
    return LBM_OK;
}

int lbm_hip_set_graph_steps(lbm_ctx *c, int graph_steps) {
    if (!c || graph_steps < 0) return fail(LBM_ERR_ARG, "bad graph length");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (graph_steps != c->graph_steps) drop_graphs(c);
    c->graph_steps = graph_steps;
    return LBM_OK;
}

int lbm_hip_query(lbm_ctx *c) {
    if (!c) return fail(LBM_ERR_ARG, "null context");
    hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return hip_fail(e, "hipStreamQuery");
}

int lbm_hip_pool_reserve(size_t bytes) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(LBM_ERR_UNSUPPORTED, "device index too large");
    std::lock_guard<std::mutex> lock(g_pool_mu);
    if (!g_pool[dev]) {
        hipMemPoolProps props = {};
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        HIPCHK(hipMemPoolCreate(&pool, &props));
        // freed blocks stay in the pool for the next context
        uint64_t keep = UINT64_MAX;
        hipError_t e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        if (e != hipSuccess) {
            (void)hipMemPoolDestroy(pool);
            return hip_fail(e, "hipMemPoolSetAttribute");
        }
        g_pool[dev] = pool;
        g_pool_users[dev] = 0;
    }
    if (bytes > 0) {
        // grow the pool once: allocate and free `bytes` on a private stream
        hipStream_t s = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        void *p = nullptr;
        hipError_t e = hipMallocFromPoolAsync(&p, bytes, g_pool[dev], s);
        if (e == hipSuccess) e = hipFreeAsync(p, s);
        hipError_t e2 = hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
        if (e != hipSuccess) return hip_fail(e, "reserving the pool");
        if (e2 != hipSuccess) return hip_fail(e2, "reserving the pool");
    }
    return LBM_OK;
}

int lbm_hip_pool_release(void) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(LBM_ERR_UNSUPPORTED, "device index too large");
    std::lock_guard<std::mutex> lock(g_pool_mu);
    if (!g_pool[dev]) return LBM_OK;
    if (g_pool_users[dev] > 0)
        return fail(LBM_ERR_STATE, std::to_string(g_pool_users[dev]) +
                                       " contexts still allocate from the pool");
    HIPCHK(hipMemPoolDestroy(g_pool[dev]));
    g_pool[dev] = nullptr;
    return LBM_OK;
}

int lbm_hip_device_memory(size_t *free_bytes, size_t *total_bytes) {
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return LBM_OK;
}

int lbm_hip_set_tile(lbm_ctx *c, int tile_x, int tile_y) {
    if (!c || tile_x < 0 || tile_y < 0) return fail(LBM_ERR_ARG, "bad tile");
    if (tile_x != c->tile_x || tile_y != c->tile_y) {
        // a replay still in flight owns its executable graph
        HIPCHK(hipStreamSynchronize(c->stream));
        drop_graphs(c);
    }
    c->tile_x = tile_x;
    c->tile_y = tile_y;
    return LBM_OK;
}

int lbm_hip_set_side_params(lbm_ctx *c, int side, const double *p0, const double *p1) {
    if (!c || side < 0 || side > 3) return fail(LBM_ERR_ARG, "bad side");
    HIPCHK(hipStreamSynchronize(c->stream));
    int n = (side == LBM_SIDE_UPPER || side == LBM_SIDE_LOWER) ? c->prob.lay.lx + 2
                                                               : c->prob.lay.ly + 2;
    if (p0) HIPCHK(hipMemcpy(c->d_p0[side], p0, n * sizeof(double), hipMemcpyHostToDevice));
    if (p1) HIPCHK(hipMemcpy(c->d_p1[side], p1, n * sizeof(double), hipMemcpyHostToDevice));
    return LBM_OK;
}

int lbm_hip_macros(lbm_ctx *c, double *rho, double *ux, double *uy) {
    if (!c) return fail(LBM_ERR_ARG, "null context");
    const Layout &L = c->prob.lay;
    // the context's own scratch, whole columns at a time (allocated on the
    // first call, counted in the footprint)
    if (!c->d_fields) {
        c->fields_nodes = fields_nodes_of(L.lx, L.ly);
        HIPCHK(ctx_malloc(c, (void **)&c->d_fields, 3 * (size_t)c->fields_nodes * sizeof(double)));
    }
    const int chunk = (int)(c->fields_nodes / L.ly);
    double *d = c->d_fields;
    const size_t cap = (size_t)c->fields_nodes;
    for (int x0 = 1; x0 <= L.lx; x0 += chunk) {
        const int nx = x0 + chunk - 1 <= L.lx ? chunk : L.lx - x0 + 1;
        const size_t m = (size_t)nx * L.ly, at = (size_t)(x0 - 1) * L.ly;
        HIPCHK(c->storage == LBM_STORE_F64
                   ? launch_macros_f64(L, x0, nx, c->collided, dbuf(c, c->cur), d, d + cap,
                                       d + 2 * cap, c->stream)
                   : launch_macros_lp(c->storage, L, x0, nx, c->collided, c->buf[c->cur], d,
                                      d + cap, d + 2 * cap, c->stream));
        double *out[3] = {rho, ux, uy};
        for (int k = 0; k < 3; ++k)
            if (out[k])
                HIPCHK(hipMemcpyAsync(out[k] + at, d + k * cap, m * sizeof(double),
                                      hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return LBM_OK;
}

int lbm_hip_diagnostics(lbm_ctx *c, double out[4]) {
    if (!c || !out) return fail(LBM_ERR_ARG, "null argument");
    if (c->storage == LBM_STORE_F64) {
        HIPCHK(launch_diag_f64(c->prob, c->collided, dbuf(c, c->cur), c->d_partial, c->d_out4,
                               c->stream));
    } else {
        HIPCHK(launch_diag_lp(c->storage, c->prob, c->collided, c->buf[c->cur], c->d_partial,
                              c->stream));
        HIPCHK(launch_diag_final(c->d_partial, c->npartial, c->d_out4, c->stream));
    }
    HIPCHK(hipMemcpyAsync(out, c->d_out4, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_hip_kernel_timing(lbm_ctx *c, int on) {
    if (!c) return fail(LBM_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->timing = on != 0;
    c->tlog.clear();
    return LBM_OK;
}

int lbm_hip_kernel_times(lbm_ctx *c, lbm_kernel_record *out, int max, int *count) {
    if (!c || !count || max < 0 || (max > 0 && !out)) return fail(LBM_ERR_ARG, "bad argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<lbm_kernel_record> recs;
    for (size_t k = 0; k < c->tlog.size(); ++k) {
        const lbm_ctx::Timed &t = c->tlog[k];
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->tev[2 * k], c->tev[2 * k + 1]));
        size_t j = 0;
        while (j < recs.size() && !(recs[j].depth == t.depth && t.name == recs[j].kernel)) ++j;
        if (j == recs.size()) {
            lbm_kernel_record r{};
            std::snprintf(r.kernel, sizeof r.kernel, "%s", t.name.c_str());
            r.depth = t.depth;
            r.nodes = t.nodes;
            r.min_ms = ms;
            r.max_ms = ms;
            recs.push_back(r);
        }
        recs[j].launches += 1;
        recs[j].ms += ms;
        recs[j].min_ms = ms < recs[j].min_ms ? ms : recs[j].min_ms;
        recs[j].max_ms = ms > recs[j].max_ms ? ms : recs[j].max_ms;
    }
    *count = (int)recs.size();
    for (int j = 0; j < max && j < (int)recs.size(); ++j) out[j] = recs[j];
    return LBM_OK;
}

int64_t lbm_hip_steps_done(const lbm_ctx *c) { return c ? c->steps : -1; }

int lbm_hip_sync(lbm_ctx *c) {
    if (!c) return fail(LBM_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_hip_time_steps(lbm_ctx *c, int variant, int nsteps, double *ms_total) {
    if (!c || !ms_total) return fail(LBM_ERR_ARG, "null argument");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    // settle the state convention outside the timed region
    int rc = LBM_OK;
    if (variant == LBM_ORIGIN) rc = to_streamed(c);
    hipError_t e = hipEventRecord(a, c->stream);
    if (!rc && e == hipSuccess) rc = step_impl(c, variant, nsteps);
    if (!rc && e == hipSuccess) e = hipEventRecord(b, c->stream);
    if (!rc && e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "event timing");
    *ms_total = ms;
    return LBM_OK;
}

double lbm_hip_bytes_per_node(const lbm_ctx *c, int variant) {
    if (!c) return 0.0;
    double pop = 9.0 * elem_size(c->storage);
    switch (variant) {
    case LBM_ORIGIN: return 4 * pop;        // collide r+w, stream r+w
    case LBM_COMBINE:
    case LBM_TIGHT: return 2 * pop;         // one read and one write per step
    case LBM_TIGHT_BLOCK:
    case LBM_PANEL: return pop;             // one read and one write per two steps
    case LBM_PANEL4: return pop / 2;        // per four steps
    case LBM_PANEL6: return pop / 3;        // per six steps
    default: return 0.0;
    }
}

size_t lbm_hip_footprint(const lbm_ctx *c) { return c ? c->footprint : 0; }

int lbm_hip_apply_node_op(int which, int n, double *f, const double *p0, const double *p1,
                          double omega) {
    if (which < 0 || which > 17 || n < 0 || (n > 0 && (!f || !p0 || !p1)))
        return fail(LBM_ERR_ARG, "bad node-operator call");
    if (n == 0) return LBM_OK;
    double *d = nullptr;
    HIPCHK(hipMalloc((void **)&d, (size_t)n * 11 * sizeof(double)));
    hipError_t e = hipMemcpy(d, f, (size_t)n * 9 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 9 * (size_t)n, p0, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 10 * (size_t)n, p1, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_node_op_f64(which, n, d, d + 9 * (size_t)n, d + 10 * (size_t)n, omega, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(f, d, (size_t)n * 9 * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "node operator");
    return LBM_OK;
}

int lbm_hip_apply_node_op_noalloc(int which, int n, double *f, const double *p0,
                                  const double *p1, double omega) {
    if (which < 0 || which > 17 || n < 0 || (n > 0 && (!f || !p0 || !p1)))
        return fail(LBM_ERR_ARG, "bad node-operator call");
    if (n > LBM_NODE_OP_MAX)
        return fail(LBM_ERR_ARG, "n = " + std::to_string(n) + " exceeds LBM_NODE_OP_MAX (" +
                                     std::to_string(LBM_NODE_OP_MAX) + ")");
    if (n == 0) return LBM_OK;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(LBM_ERR_UNSUPPORTED, "device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu[dev]);
    if (!g_scratch_ptr[dev]) {
        // first call on this device, with it current: its copy's address
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double *>(ptr);
    }
    double *d = g_scratch_ptr[dev];
    const size_t nn = (size_t)n;
    hipStream_t st = nullptr;
    HIPCHK(hipMemcpyAsync(d, f, nn * 9 * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + 9 * nn, p0, nn * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d + 10 * nn, p1, nn * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(launch_node_op_f64(which, n, d, d + 9 * nn, d + 10 * nn, omega, st));
    HIPCHK(hipMemcpyAsync(f, d, nn * 9 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return LBM_OK;
}

int lbm_hip_copy_bandwidth(size_t bytes, int iters, double *ms_per_copy) {
    if (bytes < 16 || iters < 1 || !ms_per_copy) return fail(LBM_ERR_ARG, "bad copy request");
    bytes &= ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&b, bytes);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipMemset(a, 0, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_copy(a, b, (int64_t)bytes, 0);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, 0);
    for (int k = 0; k < iters && e == hipSuccess; ++k) e = launch_copy(a, b, (int64_t)bytes, 0);
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return hip_fail(e, "copy bandwidth");
    *ms_per_copy = ms / iters;
    return LBM_OK;
}

}  // extern "C"
