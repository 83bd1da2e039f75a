// sph_api.hip - C ABI of the solver (include/sph.h): buffer ownership, the
// three-pass surface, graph-captured steps, diagnostics and kernel timing.
//
// Host counterpart in the reference: NPR-SPH/Main.cpp (init_particles
// :510-539, sendUniforms :254-282, display :291-304, reload_shader
// :428-451) and InitShader.cpp:47-112 for the error convention.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sph.h"
#include "sph_internal.h"

using namespace nprsph;

namespace {

thread_local std::string g_create_error;

constexpr double kPi = (double)3.141592741f;  // rho_pres_comp.glsl:8
constexpr int kUnroll = 16;                    // steps per captured graph
// sph_step keeps at most this many graph launches queued on the stream: the
// next launch first waits for the one kMaxGraphsInFlight launches before it.
// That still leaves 7 x 16 steps of work queued ahead of the GPU.
constexpr int kMaxGraphsInFlight = 8;

struct GridGeom {
    float lo[3];
    float inv_cs;
    double cs;
    int32_t dims[3];
    uint32_t ncells;
};

// Cell edge cs = h*(1+2^-10)/sub along y and z and cs/kXSplit along x: with
// float32 rounding of the hash, |xi-xj| < h still lands within `sub` cells
// along y and z and kXSplit*sub along x.  Grown by 2^(1/4) until each axis has
// <= 2048 cells and the grid, z rounded up to whole bands of kCellBand planes,
// has <= min(kXSplit * max(2^19, 16n), 2^30) cells (global cell indices stay
// below 2^31 for a batch of a few scenes of that size).
bool grid_geometry(float h, const sph_boundary_t& b, uint64_t n, int sub, GridGeom& g)
{
    if (!(h > 0.0f) || !std::isfinite(h) || sub < 1 || sub > 3) return false;
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(b.lower[a]) || !std::isfinite(b.upper[a])) return false;
    double cs = (double)h * (1.0 + 1.0 / 1024.0) / sub;
    const uint64_t cap = std::min<uint64_t>((uint64_t)kXSplit * std::max<uint64_t>(524288u, 16u * n), 1ull << 30);
    for (;;) {
        uint64_t total = 1;
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            const double ext = (double)b.upper[a] - (double)b.lower[a];
            double d = std::ceil(ext / (a == 0 ? cs / kXSplit : cs));
            if (!(d >= 1.0)) d = 1.0;
            if (d > 2048.0) {
                ok = false;
                d = 2048.0;
            }
            g.dims[a] = (int32_t)d;
            total *= a == 2 ? (uint64_t)((g.dims[a] + kCellBand - 1) / kCellBand * kCellBand) : (uint64_t)g.dims[a];
        }
        if (ok && total <= cap) {
            g.ncells = (uint32_t)total;
            break;
        }
        cs *= 1.189207115002721;  // 2^(1/4)
    }
    g.cs = cs;
    g.inv_cs = (float)(1.0 / cs);
    for (int a = 0; a < 3; ++a) g.lo[a] = b.lower[a];
    return true;
}

bool finite_all(const float* v, int n)
{
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

struct PendingTimer {
    const char* name;
    hipEvent_t e0, e1;
};

// Scratch of the allocation-free entry point sph_metrics_noalloc: per-tile
// partial sums of up to kNoAllocTiles tiles (kNoAllocTiles * kTile =
// 4,194,304 particles per scene), then the 7 combined values.
constexpr int kNoAllocTiles = 16384;
constexpr size_t kScratchDoubles = (size_t)kNoAllocTiles * 8u + 8u;

}  // namespace

// Every GPU the library runs on holds its own copy of this array (1 MiB) until
// the process ends.  Host code never passes the name itself to a kernel or a
// copy: on the host it is only a stand-in, and the device address comes from
// hipGetSymbolAddress (resolve_scratch).
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be the current device;
// looked up once per device.
hipError_t resolve_scratch(int device, double** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

struct sph_ctx {
    int device = 0;
    double* scratch = nullptr;  // device address of `scratch` on `device`
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int nscenes = 0;
    uint32_t n = 0;
    std::vector<uint32_t> first, count;
    std::vector<int> tile_begin, tile_end;
    std::vector<sph_constants_t> C;
    std::vector<sph_boundary_t> B;
    std::vector<sph_fixed_t> F;
    std::vector<GridGeom> geom;
    std::vector<SceneDev> sd;
    std::vector<Tile> tiles_h;      // kTile-particle tiles
    std::vector<Tile> tiles128_h;   // 128-particle tiles (gathers)
    std::vector<int> gtile_begin[2], gtile_end[2];
    uint32_t total_cells = 0;
    uint32_t cell_cap = 0;  // capacity of count/start (start has cell_cap+1)
    int bsum_cap = 0;

    DeviceState d{};
    SceneDev* d_scenes = nullptr;
    Tile* d_tiles = nullptr;
    Tile* d_tiles128 = nullptr;
    sph_particle_t* d_backup = nullptr;  // autotune snapshot of the AoS buffer

    bool b_valid = false;    // B pos/vel == AoS pos/vel through perm
    bool a_valid = false;    // A pos/vel == AoS through perm and the grid matches A
    bool rho_valid = false;  // rho, PA.w, VA.w (A order) == AoS extras through perm
    // B's cells are counted (count[], key, slot): the next build starts at the
    // scan.  The fused force+integrate kernel leaves this state.
    bool counted = false;
    // count[] may hold counts no build will consume (the state was dropped or
    // the grid changed after a counting step): cleared before the next count.
    bool counts_dirty = false;
    // bsum holds the scan's block sums of count[]: both cleared by a build,
    // then counted together by the fused step.  A count by k_hash_count
    // leaves bsum behind, and the next build first runs k_scan_reduce.
    bool bsum_valid = true;

    hipGraphExec_t graph[2] = {nullptr, nullptr};  // 1 step, kUnroll steps
    hipEvent_t inflight[kMaxGraphsInFlight] = {};  // completion of recent graph launches
    uint64_t graph_launches = 0;                    // graph launches so far (indexes inflight)

    MlpWork mlp;                               // learned coarse step
    bool mlp_loaded = false;
    int64_t adam_step = 0;
    sph_particle_t* d_judge_s = nullptr;       // judge: state before the coarse step
    sph_particle_t* d_judge_p = nullptr;       // judge: state after 4 plain steps
    float* d_pred_aos = nullptr;               // sph_mlp_predict output staging

    bool timing = false;
    std::vector<PendingTimer> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<double, int64_t>> times;

    std::string err;
};

namespace {

int fail(sph_ctx* c, int code, const std::string& msg)
{
    c->err = msg;
    return code;
}

int hip_fail(sph_ctx* c, hipError_t e, const char* what)
{
    (void)hipGetLastError();
    return fail(c, SPH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define SPH_HIP(ctx, call)                                  \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

template <class T>
hipError_t dalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void**)p, count * sizeof(T));
}

SceneDev make_scene(const sph_constants_t& c, const sph_boundary_t& b, const sph_fixed_t& f, const GridGeom& g,
                    uint32_t cell_base, uint32_t p_begin, uint32_t p_end)
{
    SceneDev s{};
    for (int a = 0; a < 3; ++a) {
        s.lo[a] = g.lo[a];
        s.dims[a] = g.dims[a];
        s.grav[a] = f.gravity[a];
        s.lower[a] = b.lower[a];
        s.upper[a] = b.upper[a];
    }
    s.inv_cs = g.inv_cs;
    s.inv_csx = (float)kXSplit * g.inv_cs;  // exact: a power of two
    s.cell_base = cell_base;
    s.ncells = g.ncells;
    s.p_begin = p_begin;
    s.p_end = p_end;
    const float h = c.smoothing_coeff * f.particle_radius;  // rho_pres_comp.glsl:40, float32
    const double hd = (double)h;
    s.h = h;
    s.h2 = h * h;
    s.poly6_m = (float)((double)c.mass * 315.0 / (64.0 * kPi * std::pow(hd, 9)));
    const double k45 = 45.0 / (kPi * std::pow(hd, 6));
    s.pres_m = (float)((double)c.mass * k45 * 0.5);
    s.visc_m = (float)((double)c.mass * k45 * (double)c.visc);
    s.rho0 = c.resting_rho;
    s.gas_const = f.gas_const;
    s.dt = f.dt;
    s.damping = f.damping;
    s.mass = c.mass;
    return s;
}

int validate_params(sph_ctx* c, const sph_constants_t& k, const sph_boundary_t& b, const sph_fixed_t& f)
{
    const float kc[4] = {k.mass, k.smoothing_coeff, k.visc, k.resting_rho};
    const float kf[7] = {f.particle_radius, f.gas_const, f.gravity[0], f.gravity[1], f.gravity[2], f.damping, f.dt};
    if (!finite_all(kc, 4) || !finite_all(kf, 7) || !finite_all(b.upper, 3) || !finite_all(b.lower, 3))
        return fail(c, SPH_ERR_INVALID, "non-finite constant, boundary or fixed parameter");
    const float h = k.smoothing_coeff * f.particle_radius;
    if (!(h > 0.0f) || !std::isfinite(h)) return fail(c, SPH_ERR_INVALID, "smoothing length must be finite and > 0");
    return SPH_OK;
}

// Graph executables may still be running on the stream (sph_step returns
// after enqueueing them, and a parameter change that moves the grid destroys
// them), so the stream is drained before they are destroyed.
void destroy_graphs(sph_ctx* c)
{
    bool any = false;
    for (auto g : c->graph) any = any || g != nullptr;
    if (any) (void)hipStreamSynchronize(c->stream);
    for (auto& g : c->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

sph_tuning_t default_tuning() { return sph_tuning_t{2, 1, 128, 0}; }

void select_gather_tiles(sph_ctx* c)
{
    if (c->d.tuning.block == 128) {
        c->d.gtiles = c->d_tiles128;
        c->d.ngtiles = (int)c->tiles128_h.size();
    } else {
        c->d.gtiles = c->d_tiles;
        c->d.ngtiles = c->d.ntiles;
    }
}

void drop_counts(sph_ctx* c);

// (Re)compute every scene's grid and parameter block; grow the cell buffers
// when needed and upload the parameter blocks.
int apply_params(sph_ctx* c, bool geometry_may_change)
{
    uint32_t cells = 0;
    std::vector<GridGeom> geom(c->nscenes);
    for (int s = 0; s < c->nscenes; ++s) {
        if (!grid_geometry(c->C[s].smoothing_coeff * c->F[s].particle_radius, c->B[s], c->count[s],
                           c->d.tuning.cells_per_h, geom[s]))
            return fail(c, SPH_ERR_INVALID, "invalid grid geometry");
        const uint64_t next = (uint64_t)cells + geom[s].ncells + 1u;  // + NaN cell
        if (next > 0x7fffffffull) return fail(c, SPH_ERR_INVALID, "too many grid cells");
        cells = (uint32_t)next;
    }
    bool changed = cells != c->total_cells;
    for (int s = 0; s < c->nscenes && !changed; ++s)
        changed = std::memcmp(&geom[s], &c->geom[s], sizeof(GridGeom)) != 0;
    (void)geometry_may_change;

    if (cells > c->cell_cap) {
        SPH_HIP(c, hipStreamSynchronize(c->stream));
        if (c->d.count) SPH_HIP(c, hipFree(c->d.count));
        if (c->d.start) SPH_HIP(c, hipFree(c->d.start));
        c->d.count = c->d.start = nullptr;
        const uint32_t cap = cells + cells / 4u;
        if (dalloc(&c->d.count, cap) != hipSuccess || dalloc(&c->d.start, (size_t)cap + 1u) != hipSuccess) {
            (void)hipGetLastError();
            c->cell_cap = 0;
            return fail(c, SPH_ERR_NOMEM, "cannot allocate grid cells");
        }
        c->cell_cap = cap;
        c->counted = false;
        c->counts_dirty = true;  // cleared (with the block sums) before the next count
    }
    const int nb = scan_blocks(cells);
    if (nb > c->bsum_cap) {
        SPH_HIP(c, hipStreamSynchronize(c->stream));
        if (c->d.bsum) SPH_HIP(c, hipFree(c->d.bsum));
        c->d.bsum = nullptr;
        if (dalloc(&c->d.bsum, (size_t)nb) != hipSuccess) {
            (void)hipGetLastError();
            c->bsum_cap = 0;
            return fail(c, SPH_ERR_NOMEM, "cannot allocate scan scratch");
        }
        c->bsum_cap = nb;
        c->counted = false;
        c->counts_dirty = true;
    }

    uint32_t base = 0;
    for (int s = 0; s < c->nscenes; ++s) {
        c->sd[s] = make_scene(c->C[s], c->B[s], c->F[s], geom[s], base, c->first[s], c->first[s] + c->count[s]);
        base += geom[s].ncells + 1u;
        hipError_t e = launch_set_scene(c->stream, c->d_scenes + s, c->sd[s]);
        if (e != hipSuccess) return hip_fail(c, e, "upload scene parameters");
    }
    c->geom = geom;
    c->total_cells = cells;
    c->d.total_cells = cells;
    if (changed) {
        destroy_graphs(c);
        c->a_valid = false;
        drop_counts(c);  // counted with the old grid
    }
    return SPH_OK;
}

hipEvent_t take_event(sph_ctx* c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}

// Enqueue one kernel through its launcher; in timing mode bracket it with
// events on the context's stream.
template <class F>
int run(sph_ctx* c, const char* name, F&& launch)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (e0) (void)hipEventRecord(e0, c->stream);
    }
    const hipError_t e = launch();
    if (c->timing && e0 && e1) {
        (void)hipEventRecord(e1, c->stream);
        c->pending.push_back({name, e0, e1});
    }
    if (e != hipSuccess) return hip_fail(c, e, name);
    return SPH_OK;
}

#define SPH_RUN(ctx, name, expr)                           \
    do {                                                   \
        int rc_ = run(ctx, name, [&]() { return (expr); }); \
        if (rc_ != SPH_OK) return rc_;                     \
    } while (0)

// Cell counts of B (valid B required), unless the last fused step left them.
int ensure_counted(sph_ctx* c)
{
    if (c->counted) return SPH_OK;
    if (c->counts_dirty) {
        SPH_RUN(c, "clear_counts", launch_fill_u32(c->stream, c->d.count, c->cell_cap, 0u));
        SPH_RUN(c, "clear_counts", launch_fill_u32(c->stream, c->d.bsum, (uint32_t)c->bsum_cap, 0u));
        c->counts_dirty = false;
    }
    SPH_RUN(c, "hash_count", launch_hash_count(c->stream, c->d));
    c->counted = true;
    c->bsum_valid = false;
    return SPH_OK;
}

// The scan's block sums of the counted B, summed from count[] unless the fused
// step counted them.
int ensure_block_sums(sph_ctx* c)
{
    if (c->bsum_valid) return SPH_OK;
    SPH_RUN(c, "scan_reduce", launch_scan_reduce(c->stream, c->d));
    c->bsum_valid = true;
    return SPH_OK;
}

// The counts drop out of use: a later count must clear them first.
void drop_counts(sph_ctx* c)
{
    if (c->counted) c->counts_dirty = true;
    c->counted = false;
}

int enqueue_build(sph_ctx* c)
{
    const DeviceState& d = c->d;
    int rc = ensure_counted(c);
    if (rc != SPH_OK) return rc;
    if ((rc = ensure_block_sums(c)) != SPH_OK) return rc;
    if (!scan_folds(d.total_cells)) SPH_RUN(c, "scan_top", launch_scan_top(c->stream, d));
    SPH_RUN(c, "scan_down", launch_scan_down(c->stream, d));  // also clears count[]
    c->counted = false;
    SPH_RUN(c, "scatter", launch_scatter(c->stream, d));  // also clears bsum
    c->bsum_valid = true;  // count[] and bsum both zero: the fused step counts both
    SPH_RUN(c, "rank_reorder", launch_rank_reorder(c->stream, d));
    return SPH_OK;
}

// One step on the sorted state.  write_aos = false leaves the AoS buffer to a
// later step of the same call: no kernel of a step reads it.  The fused kernel
// also counts the new state into the next grid's cells.
int enqueue_step(sph_ctx* c, bool write_aos)
{
    int rc = enqueue_build(c);
    if (rc != SPH_OK) return rc;
    SPH_RUN(c, "density", launch_density(c->stream, c->d, false));
    SPH_RUN(c, "force_integrate",
            launch_force(c->stream, c->d, write_aos ? ForceMode::FusedAos : ForceMode::Fused));
    c->counted = true;
    return SPH_OK;
}

int ensure_b(sph_ctx* c)
{
    if (!c->b_valid) {
        SPH_RUN(c, "refresh", launch_refresh(c->stream, c->d, true, false));
        c->b_valid = true;
    }
    return SPH_OK;
}

// Captures `steps` steps.  Every graph starts and ends with B counted and the
// scan's block sums valid (the caller sees to both before the first launch),
// so its first build has no hash and no reduction; capturing launches
// nothing, so the flags are restored afterwards.
int capture_graph(sph_ctx* c, int which, int steps)
{
    hipGraph_t g = nullptr;
    const bool counted = c->counted, dirty = c->counts_dirty, bsum_valid = c->bsum_valid;
    c->counted = true;
    c->bsum_valid = true;
    SPH_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = SPH_OK;
    for (int k = 0; k < steps && rc == SPH_OK; ++k) rc = enqueue_step(c, k == steps - 1);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->counted = counted;
    c->counts_dirty = dirty;
    c->bsum_valid = bsum_valid;
    if (rc != SPH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&c->graph[which], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        c->graph[which] = nullptr;
        return hip_fail(c, e, "hipGraphInstantiate");
    }
    return SPH_OK;
}

void invalidate_state(sph_ctx* c)
{
    c->a_valid = c->b_valid = c->rho_valid = false;
    drop_counts(c);
}

int check_scene(sph_ctx* c, int scene)
{
    if (scene < -1 || scene >= c->nscenes) return fail(c, SPH_ERR_INVALID, "scene index out of range");
    return SPH_OK;
}

int set_device(sph_ctx* c)
{
    SPH_HIP(c, hipSetDevice(c->device));
    return SPH_OK;
}

void free_all(sph_ctx* c)
{
    destroy_graphs(c);
    for (auto& p : c->pending) {
        (void)hipEventDestroy(p.e0);
        (void)hipEventDestroy(p.e1);
    }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    for (auto e : c->inflight)
        if (e) (void)hipEventDestroy(e);
    const MlpWork& w = c->mlp;
    void* bufs[] = {c->d.aos, c->d.PA, c->d.VA, c->d.PB, c->d.VB, c->d.rho, c->d.perm, c->d.key,
                    c->d.slot, c->d.torig, c->d.tsrc, c->d.tkey, c->d.count, c->d.start, c->d.bsum, c->d.partials,
                    c->d_scenes, c->d_tiles, c->d_tiles128, c->d_backup, c->d_judge_s, c->d_judge_p, c->d_pred_aos,
                    w.params, w.adam_m, w.adam_v, w.grad, w.cnt, w.off, w.tgt, w.wt, w.pred, w.dpred, w.lossp,
                    w.pj, w.pt, w.X, w.Z1, w.Z2, w.dZ1, w.dZ2, w.M, w.partial, w.pack_a, w.pack_c, w.pairs, w.gelu_tab};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
}

// ---- learned coarse step helpers ----

template <class T>
int grow(sph_ctx* c, T** p, uint64_t& cap, uint64_t need, uint64_t per_item, const char* what)
{
    if (need <= cap && *p) return SPH_OK;
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    if (*p) SPH_HIP(c, hipFree(*p));
    *p = nullptr;
    const uint64_t want = need + need / 4 + 1;
    if (dalloc(p, (size_t)(want * per_item)) != hipSuccess) {
        (void)hipGetLastError();
        cap = 0;
        return fail(c, SPH_ERR_NOMEM, std::string("cannot allocate ") + what);
    }
    cap = want;
    return SPH_OK;
}

int ensure_mlp_base(sph_ctx* c)
{
    MlpWork& w = c->mlp;
    if (w.params) return SPH_OK;
    if (dalloc(&w.params, kMlpParams) != hipSuccess || dalloc(&w.adam_m, kMlpParams) != hipSuccess ||
        dalloc(&w.adam_v, kMlpParams) != hipSuccess || dalloc(&w.grad, kMlpParams) != hipSuccess ||
        dalloc(&w.pack_a, kMlpPackA) != hipSuccess || dalloc(&w.pack_c, kMlpPackC) != hipSuccess ||
        dalloc(&w.gelu_tab, kMlpGeluTab) != hipSuccess ||
        dalloc(&w.pairs, 1) != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, SPH_ERR_NOMEM, "cannot allocate network parameters");
    }
    SPH_HIP(c, hipMemsetAsync(w.adam_m, 0, sizeof(float) * kMlpParams, c->stream));
    SPH_HIP(c, hipMemsetAsync(w.adam_v, 0, sizeof(float) * kMlpParams, c->stream));
    return SPH_OK;
}

// Per-target buffers sized for T targets.
int ensure_targets(sph_ctx* c, uint64_t T)
{
    MlpWork& w = c->mlp;
    if (T <= w.target_cap && w.cnt) return SPH_OK;
    uint64_t cap = w.target_cap;
    int rc;
    uint64_t c1 = cap, c2 = cap, c3 = cap, c4 = cap, c5 = cap, c6 = cap, c7 = cap;
    // Until every buffer has grown, the recorded capacity is 0, so a failed
    // grow (which leaves its buffer null) makes the next call regrow them all.
    w.target_cap = 0;
    if ((rc = grow(c, &w.cnt, c1, T, 1, "pair counts")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.off, c2, T + 1, 1, "pair offsets")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.tgt, c3, T, 3, "targets")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.wt, c4, T, 1, "target weights")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.pred, c5, T, 3, "predictions")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.dpred, c6, T, 3, "prediction gradients")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.lossp, c7, T, 1, "loss terms")) != SPH_OK) return rc;
    w.target_cap = std::min({c1, c2 - 1, c3, c4, c5, c6, c7});
    return SPH_OK;
}

int ensure_pairs(sph_ctx* c, uint64_t P)
{
    MlpWork& w = c->mlp;
    if (P <= w.pair_cap && w.pj) return SPH_OK;
    uint64_t cap[10];
    for (auto& x : cap) x = w.pair_cap;
    w.pair_cap = 0;  // as in ensure_targets: a failed grow forces a full regrow next time
    int rc;
    if ((rc = grow(c, &w.pj, cap[0], P, 1, "pairs")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.pt, cap[1], P, 1, "pair targets")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.X, cap[2], P, kMlpIn, "pair features")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.Z1, cap[3], P, kMlpHidden, "layer-1 activations")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.Z2, cap[4], P, kMlpHidden, "layer-2 activations")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.dZ1, cap[5], P, kMlpHidden, "layer-1 gradients")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.dZ2, cap[6], P, kMlpHidden, "layer-2 gradients")) != SPH_OK) return rc;
    if ((rc = grow(c, &w.M, cap[7], P, kMlpOut, "messages")) != SPH_OK) return rc;
    w.pair_cap = *std::min_element(cap, cap + 8);
    return SPH_OK;
}

// Density pass on the current AoS state (sph_rho_pres without the public
// entry point), leaving A, rho, PA.w/VA.w and the grid valid.
int density_pass(sph_ctx* c, bool write_aos)
{
    int rc;
    if ((rc = ensure_b(c)) != SPH_OK) return rc;
    if ((rc = enqueue_build(c)) != SPH_OK) return rc;
    c->b_valid = false;
    c->a_valid = true;
    SPH_RUN(c, "density", launch_density(c->stream, c->d, write_aos));
    c->rho_valid = true;
    return SPH_OK;
}

int mean_rho(sph_ctx* c, double* out);

// Per-scene metrics from the combined per-tile sums acc[0..6].
void finish_metrics(const double* acc, uint32_t count, sph_metrics_t* out)
{
    const double cnt = (double)count;
    const double fin = std::max(cnt - acc[6], 1.0);
    out->mean_rho = acc[0] / cnt;
    out->kinetic = acc[1];
    out->com[0] = acc[2] / fin;
    out->com[1] = acc[3] / fin;
    out->com[2] = acc[4] / fin;
    out->max_speed = acc[5];
    out->nonfinite = (uint64_t)acc[6];
}

}  // namespace

extern "C" {

void sph_defaults(sph_constants_t* c, sph_boundary_t* b, sph_fixed_t* f)
{
    if (c) *c = sph_constants_t{0.02f, 4.0f, 3000.0f, 1000.0f};  // Main.cpp:110-116
    if (b) *b = sph_boundary_t{{0.5f, 1.0f, 0.5f, 1.0f}, {-0.1f, -0.35f, -0.1f, 1.0f}};  // Main.cpp:118-122
    if (f) {
        f->particle_radius = 0.005f;  // rho_pres_comp.glsl:5
        f->gas_const = 2000.0f;       // rho_pres_comp.glsl:33
        f->gravity[0] = 0.0f;         // force_comp.glsl:33
        f->gravity[1] = -9806.65f;
        f->gravity[2] = 0.0f;
        f->damping = 0.3f;              // integrate_comp.glsl:8
        f->dt = 1.0f / 10000.0f;        // integrate_comp.glsl:33 at NUM_PARTICLES = 10000
    }
}

const char* sph_create_error(void) { return g_create_error.c_str(); }

int sph_create_batch(sph_ctx** out, int nscenes, const size_t* counts, const sph_constants_t* cs,
                     const sph_boundary_t* bs, const sph_fixed_t* fs, void* stream)
{
    g_create_error.clear();
    if (!out) {
        g_create_error = "out is null";
        return SPH_ERR_INVALID;
    }
    *out = nullptr;
    if (nscenes < 1 || !counts) {
        g_create_error = "need at least one scene";
        return SPH_ERR_INVALID;
    }
    uint64_t total = 0;
    for (int s = 0; s < nscenes; ++s) {
        if (counts[s] < 1) {
            g_create_error = "every scene needs at least one particle";
            return SPH_ERR_INVALID;
        }
        total += counts[s];
    }
    if (total > (1ull << 28) - 2u) {  // slot arrays (plus one sentinel) stay under 2^32 bytes of float4
        g_create_error = "too many particles (limit 2^28-2)";
        return SPH_ERR_INVALID;
    }
    sph_ctx* c = new sph_ctx();
    auto bail = [&](int code) {
        g_create_error = c->err;
        free_all(c);
        delete c;
        return code;
    };
    if (hipGetDevice(&c->device) != hipSuccess) {
        c->err = "no HIP device";
        (void)hipGetLastError();
        return bail(SPH_ERR_HIP);
    }
    if (const hipError_t e = resolve_scratch(c->device, &c->scratch); e != hipSuccess) {
        c->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        (void)hipGetLastError();
        return bail(SPH_ERR_HIP);
    }
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            c->err = "hipStreamCreate failed";
            (void)hipGetLastError();
            return bail(SPH_ERR_HIP);
        }
        c->own_stream = true;
    }
    c->nscenes = nscenes;
    c->n = (uint32_t)total;
    c->first.resize(nscenes);
    c->count.resize(nscenes);
    c->C.resize(nscenes);
    c->B.resize(nscenes);
    c->F.resize(nscenes);
    c->geom.resize(nscenes);
    c->sd.resize(nscenes);
    c->tile_begin.resize(nscenes);
    c->tile_end.resize(nscenes);
    uint32_t off = 0;
    for (int s = 0; s < nscenes; ++s) {
        c->first[s] = off;
        c->count[s] = (uint32_t)counts[s];
        sph_defaults(&c->C[s], &c->B[s], &c->F[s]);
        if (cs) c->C[s] = cs[s];
        if (bs) c->B[s] = bs[s];
        if (fs) c->F[s] = fs[s];
        int rc = validate_params(c, c->C[s], c->B[s], c->F[s]);
        if (rc != SPH_OK) return bail(rc);
        c->tile_begin[s] = (int)c->tiles_h.size();
        for (uint32_t b = 0; b < c->count[s]; b += kTile)
            c->tiles_h.push_back(Tile{(uint32_t)s, off + b, off + std::min<uint32_t>(b + kTile, c->count[s]), 0u});
        c->tile_end[s] = (int)c->tiles_h.size();
        for (uint32_t b = 0; b < c->count[s]; b += 128u)
            c->tiles128_h.push_back(Tile{(uint32_t)s, off + b, off + std::min<uint32_t>(b + 128u, c->count[s]), 0u});
        off += c->count[s];
    }
    const uint32_t n = c->n;
    DeviceState& d = c->d;
    d.n = n;
    d.ntiles = (int)c->tiles_h.size();
    // PA and VA carry one extra slot, the force pass's inert sentinel particle.
    bool ok = dalloc(&d.aos, n) == hipSuccess && dalloc(&d.PA, (size_t)n + 1u) == hipSuccess &&
              dalloc(&d.VA, (size_t)n + 1u) == hipSuccess &&
              dalloc(&d.PB, n) == hipSuccess && dalloc(&d.VB, n) == hipSuccess && dalloc(&d.rho, n) == hipSuccess &&
              dalloc(&d.perm, n) == hipSuccess && dalloc(&d.key, n) == hipSuccess &&
              dalloc(&d.slot, n) == hipSuccess && dalloc(&d.torig, n) == hipSuccess &&
              dalloc(&d.tsrc, n) == hipSuccess && dalloc(&d.tkey, n) == hipSuccess &&
              dalloc(&d.partials, (size_t)d.ntiles * 8u) == hipSuccess &&
              dalloc(&c->d_scenes, (size_t)nscenes) == hipSuccess && dalloc(&c->d_tiles, (size_t)d.ntiles) == hipSuccess &&
              dalloc(&c->d_tiles128, c->tiles128_h.size()) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        c->err = "device allocation failed";
        return bail(SPH_ERR_NOMEM);
    }
    d.scenes = c->d_scenes;
    d.tiles = c->d_tiles;
    d.tuning = default_tuning();
    select_gather_tiles(c);
    if (hipMemcpyAsync(c->d_tiles, c->tiles_h.data(), sizeof(Tile) * c->tiles_h.size(), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess ||
        hipMemcpyAsync(c->d_tiles128, c->tiles128_h.data(), sizeof(Tile) * c->tiles128_h.size(), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess) {
        (void)hipGetLastError();
        c->err = "tile upload failed";
        return bail(SPH_ERR_HIP);
    }
    int rc = apply_params(c, true);
    if (rc != SPH_OK) return bail(rc);
    // Sentinel: far from every particle (finite distances), p = 0, 1/rho = 0.
    static const float4 kSentinelP = make_float4(1e10f, 1e10f, 1e10f, 0.f);
    static const float4 kSentinelV = make_float4(0.f, 0.f, 0.f, 0.f);
    if (hipMemcpyAsync(d.PA + n, &kSentinelP, sizeof(float4), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(d.VA + n, &kSentinelV, sizeof(float4), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        launch_fill_aos_defaults(c->stream, d.aos, n) != hipSuccess || launch_identity(c->stream, d.perm, n) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        (void)hipGetLastError();
        c->err = "initialisation kernels failed";
        return bail(SPH_ERR_HIP);
    }
    *out = c;
    return SPH_OK;
}

int sph_create(sph_ctx** out, size_t n, const sph_constants_t* c, const sph_boundary_t* b, const sph_fixed_t* f,
               void* stream)
{
    return sph_create_batch(out, 1, &n, c, b, f, stream);
}

void sph_destroy(sph_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
}

const char* sph_last_error(const sph_ctx* c) { return c ? c->err.c_str() : "null context"; }
size_t sph_size(const sph_ctx* c) { return c ? c->n : 0; }
int sph_num_scenes(const sph_ctx* c) { return c ? c->nscenes : 0; }
void* sph_stream(sph_ctx* c) { return c ? (void*)c->stream : nullptr; }

static int set_param_common(sph_ctx* c, int scene, const sph_constants_t* k, const sph_boundary_t* b,
                            const sph_fixed_t* f)
{
    if (!c) return SPH_ERR_INVALID;
    int rc = check_scene(c, scene);
    if (rc != SPH_OK) return rc;
    rc = set_device(c);
    if (rc != SPH_OK) return rc;
    const int s0 = scene < 0 ? 0 : scene, s1 = scene < 0 ? c->nscenes : scene + 1;
    for (int s = s0; s < s1; ++s) {
        rc = validate_params(c, k ? *k : c->C[s], b ? *b : c->B[s], f ? *f : c->F[s]);
        if (rc != SPH_OK) return rc;
    }
    auto oldC = c->C;
    auto oldB = c->B;
    auto oldF = c->F;
    for (int s = s0; s < s1; ++s) {
        if (k) c->C[s] = *k;
        if (b) c->B[s] = *b;
        if (f) c->F[s] = *f;
    }
    rc = apply_params(c, true);
    if (rc != SPH_OK) {
        c->C = oldC;
        c->B = oldB;
        c->F = oldF;
    }
    return rc;
}

int sph_set_constants(sph_ctx* c, int scene, const sph_constants_t* k)
{
    if (!k) return c ? fail(c, SPH_ERR_INVALID, "constants is null") : SPH_ERR_INVALID;
    return set_param_common(c, scene, k, nullptr, nullptr);
}

int sph_set_boundary(sph_ctx* c, int scene, const sph_boundary_t* b)
{
    if (!b) return c ? fail(c, SPH_ERR_INVALID, "boundary is null") : SPH_ERR_INVALID;
    return set_param_common(c, scene, nullptr, b, nullptr);
}

int sph_set_fixed(sph_ctx* c, int scene, const sph_fixed_t* f)
{
    if (!f) return c ? fail(c, SPH_ERR_INVALID, "fixed is null") : SPH_ERR_INVALID;
    return set_param_common(c, scene, nullptr, nullptr, f);
}

int sph_get_params(const sph_ctx* c, int scene, sph_constants_t* k, sph_boundary_t* b, sph_fixed_t* f)
{
    if (!c || scene < 0 || scene >= c->nscenes) return SPH_ERR_INVALID;
    if (k) *k = c->C[scene];
    if (b) *b = c->B[scene];
    if (f) *f = c->F[scene];
    return SPH_OK;
}

int sph_upload(sph_ctx* c, const sph_particle_t* host, size_t first, size_t n)
{
    if (!c) return SPH_ERR_INVALID;
    if (!host && n) return fail(c, SPH_ERR_INVALID, "host pointer is null");
    if (first > c->n || n > c->n - first) return fail(c, SPH_ERR_INVALID, "range exceeds particle count");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    if (n) SPH_HIP(c, hipMemcpyAsync(c->d.aos + first, host, n * sizeof(sph_particle_t), hipMemcpyHostToDevice, c->stream));
    invalidate_state(c);
    return SPH_OK;
}

int sph_download(sph_ctx* c, sph_particle_t* host, size_t first, size_t n)
{
    if (!c) return SPH_ERR_INVALID;
    if (!host && n) return fail(c, SPH_ERR_INVALID, "host pointer is null");
    if (first > c->n || n > c->n - first) return fail(c, SPH_ERR_INVALID, "range exceeds particle count");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    if (n) SPH_HIP(c, hipMemcpyAsync(host, c->d.aos + first, n * sizeof(sph_particle_t), hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    return SPH_OK;
}

namespace {

constexpr char kSnapMagic[8] = {'N', 'P', 'R', 'S', 'P', 'H', 'v', '1'};

struct SnapHeader {
    char magic[8];
    uint32_t header_bytes;
    uint32_t record_bytes;
    uint64_t count;
    uint32_t nscenes;
    uint32_t reserved0;
    uint64_t frame;
    uint8_t reserved[24];
};
static_assert(sizeof(SnapHeader) == 64, "snapshot header layout");

}  // namespace

int sph_save(sph_ctx* c, const char* path, uint64_t frame)
{
    if (!c) return SPH_ERR_INVALID;
    if (!path) return fail(c, SPH_ERR_INVALID, "path is null");
    std::vector<sph_particle_t> host(c->n);
    int rc = sph_download(c, host.data(), 0, c->n);
    if (rc != SPH_OK) return rc;
    SnapHeader h{};
    std::memcpy(h.magic, kSnapMagic, sizeof(kSnapMagic));
    h.header_bytes = sizeof(SnapHeader);
    h.record_bytes = sizeof(sph_particle_t);
    h.count = c->n;
    h.nscenes = (uint32_t)c->nscenes;
    h.frame = frame;
    std::vector<uint64_t> counts(c->count.begin(), c->count.end());
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(c, SPH_ERR_INVALID, std::string("cannot open for writing: ") + path);
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 &&
              std::fwrite(counts.data(), sizeof(uint64_t), counts.size(), f) == counts.size() &&
              std::fwrite(host.data(), sizeof(sph_particle_t), host.size(), f) == host.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(c, SPH_ERR_INVALID, std::string("write failed: ") + path);
    return SPH_OK;
}

int sph_load(sph_ctx* c, const char* path, uint64_t* frame)
{
    if (!c) return SPH_ERR_INVALID;
    if (!path) return fail(c, SPH_ERR_INVALID, "path is null");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(c, SPH_ERR_INVALID, std::string("cannot open: ") + path);
    SnapHeader h{};
    std::vector<uint64_t> counts;
    std::vector<sph_particle_t> host;
    const char* why = nullptr;
    if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kSnapMagic, sizeof(kSnapMagic)) != 0)
        why = "not a snapshot file";
    else if (h.header_bytes != sizeof(SnapHeader) || h.record_bytes != sizeof(sph_particle_t))
        why = "unsupported snapshot layout";
    else if (h.count != c->n || h.nscenes != (uint32_t)c->nscenes)
        why = "snapshot particle or scene count differs from the context";
    if (!why) {
        counts.resize(h.nscenes);
        if (std::fread(counts.data(), sizeof(uint64_t), counts.size(), f) != counts.size())
            why = "truncated snapshot";
        else
            for (int s = 0; s < c->nscenes; ++s)
                if (counts[s] != (uint64_t)c->count[s]) why = "snapshot scene counts differ from the context";
    }
    if (!why) {
        host.resize(c->n);
        if (std::fread(host.data(), sizeof(sph_particle_t), host.size(), f) != host.size()) why = "truncated snapshot";
    }
    std::fclose(f);
    if (why) return fail(c, SPH_ERR_INVALID, why);
    int rc = sph_upload(c, host.data(), 0, c->n);
    if (rc != SPH_OK) return rc;
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    if (frame) *frame = h.frame;
    return SPH_OK;
}

int sph_init_lattice(sph_ctx* c, int scene, const sph_lattice_t* spec)
{
    if (!c) return SPH_ERR_INVALID;
    if (!spec) return fail(c, SPH_ERR_INVALID, "lattice spec is null");
    if (scene < 0 || scene >= c->nscenes) return fail(c, SPH_ERR_INVALID, "scene index out of range");
    if (spec->nx < 1 || spec->ny < 1 || spec->nz < 1) return fail(c, SPH_ERR_INVALID, "lattice dims must be >= 1");
    const uint64_t sites = (uint64_t)spec->nx * (uint64_t)spec->ny * (uint64_t)spec->nz;
    if (spec->count != c->count[scene] || (uint64_t)spec->count > sites)
        return fail(c, SPH_ERR_INVALID, "lattice count must equal the scene's particle count and fit the lattice");
    if (!finite_all(spec->origin, 3) || !std::isfinite(spec->spacing) || !std::isfinite(spec->jitter))
        return fail(c, SPH_ERR_INVALID, "non-finite lattice parameter");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    SPH_RUN(c, "lattice", launch_lattice(c->stream, c->d.aos, c->first[scene], *spec));
    invalidate_state(c);
    return SPH_OK;
}

// rho_pres_comp.glsl:35-58: rebuild the neighbour grid from the AoS positions,
// then the density gather; writes extras.x = rho and extras.y = p.
int sph_rho_pres(sph_ctx* c)
{
    if (!c) return SPH_ERR_INVALID;
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    if ((rc = ensure_b(c)) != SPH_OK) return rc;
    if ((rc = enqueue_build(c)) != SPH_OK) return rc;
    c->b_valid = false;
    c->a_valid = true;
    SPH_RUN(c, "density", launch_density(c->stream, c->d, true));
    c->rho_valid = true;
    return SPH_OK;
}

// force_comp.glsl:35-66: reads pos, vel and extras.xy; writes force.xyz.
int sph_force(sph_ctx* c)
{
    if (!c) return SPH_ERR_INVALID;
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    if (!c->a_valid) {
        if ((rc = ensure_b(c)) != SPH_OK) return rc;
        if ((rc = enqueue_build(c)) != SPH_OK) return rc;
        c->b_valid = false;
        c->a_valid = true;
        c->rho_valid = false;
    }
    if (!c->rho_valid) {
        SPH_RUN(c, "refresh_extras", launch_refresh(c->stream, c->d, false, true));
        c->rho_valid = true;
    }
    SPH_RUN(c, "force", launch_force(c->stream, c->d, ForceMode::Pass));
    return SPH_OK;
}

// integrate_comp.glsl:35-81 on the AoS buffer.
int sph_integrate(sph_ctx* c)
{
    if (!c) return SPH_ERR_INVALID;
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    SPH_RUN(c, "integrate", launch_integrate_aos(c->stream, c->d));
    invalidate_state(c);
    return SPH_OK;
}

int sph_step(sph_ctx* c, int nsteps)
{
    if (!c) return SPH_ERR_INVALID;
    if (nsteps < 0) return fail(c, SPH_ERR_INVALID, "nsteps must be >= 0");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    if ((rc = ensure_b(c)) != SPH_OK) return rc;
    if (!c->timing && ((rc = ensure_counted(c)) != SPH_OK || (rc = ensure_block_sums(c)) != SPH_OK))
        return rc;  // graphs start from counted B with its block sums
    if (nsteps == 0) {
        // Preparation only: the sorted, counted state (above) and both step graphs.
        if (!c->graph[1] && (rc = capture_graph(c, 1, kUnroll)) != SPH_OK) return rc;
        if (!c->graph[0] && (rc = capture_graph(c, 0, 1)) != SPH_OK) return rc;
        return SPH_OK;
    }
    auto run_steps = [&]() -> int {
        int r;
        if (c->timing) {
            for (int k = 0; k < nsteps; ++k)
                if ((r = enqueue_step(c, k == nsteps - 1)) != SPH_OK) return r;
            return SPH_OK;
        }
        const int big = nsteps / kUnroll, small = nsteps % kUnroll;
        if (big && !c->graph[1] && (r = capture_graph(c, 1, kUnroll)) != SPH_OK) return r;
        if (small && !c->graph[0] && (r = capture_graph(c, 0, 1)) != SPH_OK) return r;
        // Launch q of the context waits for launch q - kMaxGraphsInFlight
        // (possibly from an earlier call) to finish.
        for (int k = 0; k < big + small; ++k) {
            const uint64_t q = c->graph_launches;
            hipEvent_t& ev = c->inflight[q % kMaxGraphsInFlight];
            if (!ev) SPH_HIP(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            if (q >= (uint64_t)kMaxGraphsInFlight) SPH_HIP(c, hipEventSynchronize(ev));
            SPH_HIP(c, hipGraphLaunch(c->graph[k < big ? 1 : 0], c->stream));
            SPH_HIP(c, hipEventRecord(ev, c->stream));
            c->graph_launches = q + 1;
        }
        return SPH_OK;
    };
    if ((rc = run_steps()) != SPH_OK) {
        // Steps already enqueued may have advanced the sorted state past the
        // AoS buffer, which holds the last step that wrote it: the next call
        // rebuilds from the buffer.
        invalidate_state(c);
        return rc;
    }
    c->b_valid = true;
    c->a_valid = false;
    c->rho_valid = false;
    c->counted = true;  // the last fused step counted B
    return SPH_OK;
}

int sph_device_buffer(sph_ctx* c, sph_particle_t** d_aos, size_t* n)
{
    if (!c) return SPH_ERR_INVALID;
    if (d_aos) *d_aos = c->d.aos;
    if (n) *n = c->n;
    return SPH_OK;
}

int sph_synchronize(sph_ctx* c)
{
    if (!c) return SPH_ERR_INVALID;
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    return SPH_OK;
}

int sph_grid_info(const sph_ctx* c, int scene, sph_grid_info_t* out)
{
    if (!c || !out || scene < 0 || scene >= c->nscenes) return SPH_ERR_INVALID;
    const GridGeom& g = c->geom[scene];
    for (int a = 0; a < 3; ++a) {
        out->lo[a] = g.lo[a];
        out->dims[a] = g.dims[a];
    }
    out->inv_cell = g.inv_cs;
    out->cell = g.cs;
    out->ncells = g.ncells;
    out->cell_base = c->sd[scene].cell_base;
    out->first = c->first[scene];
    out->count = c->count[scene];
    out->cells_per_h = c->d.tuning.cells_per_h;
    return SPH_OK;
}

size_t sph_total_cells(const sph_ctx* c) { return c ? c->total_cells : 0; }

int sph_download_grid(sph_ctx* c, uint32_t* cell_start, size_t ncell_entries, uint32_t* perm, size_t n)
{
    if (!c) return SPH_ERR_INVALID;
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    if (cell_start) {
        if (ncell_entries != (size_t)c->total_cells + 1u) return fail(c, SPH_ERR_INVALID, "cell_start needs total_cells+1 entries");
        SPH_HIP(c, hipMemcpyAsync(cell_start, c->d.start, ncell_entries * 4u, hipMemcpyDeviceToHost, c->stream));
    }
    if (perm) {
        if (n != c->n) return fail(c, SPH_ERR_INVALID, "perm needs one entry per particle");
        SPH_HIP(c, hipMemcpyAsync(perm, c->d.perm, n * 4u, hipMemcpyDeviceToHost, c->stream));
    }
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    return SPH_OK;
}

int sph_metrics(sph_ctx* c, int scene, sph_metrics_t* out)
{
    if (!c) return SPH_ERR_INVALID;
    if (!out || scene < 0 || scene >= c->nscenes) return fail(c, SPH_ERR_INVALID, "bad metrics arguments");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    const int t0 = c->tile_begin[scene], nt = c->tile_end[scene] - t0;
    SPH_RUN(c, "metrics", launch_metrics(c->stream, c->d, t0, nt, c->d.partials));
    std::vector<double> part((size_t)nt * 8u);
    SPH_HIP(c, hipMemcpyAsync(part.data(), c->d.partials, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < 7; ++k) {
            const double x = part[(size_t)t * 8 + k];
            acc[k] = (k == 5) ? (x > acc[k] ? x : acc[k]) : acc[k] + x;
        }
    finish_metrics(acc, c->count[scene], out);
    return SPH_OK;
}

int sph_metrics_noalloc(sph_ctx* c, int scene, sph_metrics_t* out)
{
    if (!c) return SPH_ERR_INVALID;
    if (!out || scene < 0 || scene >= c->nscenes) return fail(c, SPH_ERR_INVALID, "bad metrics arguments");
    const int t0 = c->tile_begin[scene], nt = c->tile_end[scene] - t0;
    if (nt > kNoAllocTiles)
        return fail(c, SPH_ERR_INVALID,
                    "sph_metrics_noalloc: scene of " + std::to_string(c->count[scene]) +
                        " particles exceeds the allocation-free limit of " +
                        std::to_string((long long)kNoAllocTiles * kTile) + " (use sph_metrics)");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    if (!c->scratch) return fail(c, SPH_ERR_STATE, "allocation-free scratch not resolved for this device");
    double* partials = c->scratch;
    double* folded = c->scratch + (size_t)kNoAllocTiles * 8u;
    // Launched directly, not through the kernel-timing wrapper, whose events
    // may be created on demand.
    if (const hipError_t e = launch_metrics(c->stream, c->d, t0, nt, partials); e != hipSuccess)
        return hip_fail(c, e, "metrics");
    if (const hipError_t e = launch_metrics_fold(c->stream, partials, nt, folded); e != hipSuccess)
        return hip_fail(c, e, "metrics_fold");
    double acc[8];
    SPH_HIP(c, hipMemcpyAsync(acc, folded, sizeof(double) * 7u, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    finish_metrics(acc, c->count[scene], out);
    return SPH_OK;
}

int sph_get_tuning(const sph_ctx* c, sph_tuning_t* out)
{
    if (!c || !out) return SPH_ERR_INVALID;
    *out = c->d.tuning;
    return SPH_OK;
}

int sph_set_tuning(sph_ctx* c, const sph_tuning_t* t)
{
    if (!c) return SPH_ERR_INVALID;
    if (!t || !tuning_supported(*t)) return fail(c, SPH_ERR_INVALID, "unsupported tuning");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    const sph_tuning_t old = c->d.tuning;
    c->d.tuning = *t;
    select_gather_tiles(c);
    rc = apply_params(c, true);
    if (rc != SPH_OK) {
        c->d.tuning = old;
        select_gather_tiles(c);
        (void)apply_params(c, true);
        return rc;
    }
    destroy_graphs(c);  // launch shapes are baked into captured graphs
    return SPH_OK;
}

int sph_autotune(sph_ctx* c, int trial_steps, sph_tuning_t* best, double* best_ms)
{
    if (!c) return SPH_ERR_INVALID;
    if (trial_steps < 1) return fail(c, SPH_ERR_INVALID, "trial_steps must be >= 1");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    if (!c->d_backup && dalloc(&c->d_backup, c->n) != hipSuccess) {
        (void)hipGetLastError();
        c->d_backup = nullptr;
        return fail(c, SPH_ERR_NOMEM, "cannot allocate the autotune snapshot");
    }
    SPH_HIP(c, hipMemcpyAsync(c->d_backup, c->d.aos, sizeof(sph_particle_t) * c->n, hipMemcpyDeviceToDevice, c->stream));
    const sph_tuning_t start_tuning = c->d.tuning;
    // Kernel timing is off during the trials and restored, with the events
    // released, on every way out.
    struct Guard {
        sph_ctx* c;
        bool was_timing;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Guard()
        {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            c->timing = was_timing;
        }
    } guard{c, c->timing};
    c->timing = false;
    SPH_HIP(c, hipEventCreate(&guard.e0));
    SPH_HIP(c, hipEventCreate(&guard.e1));
    hipEvent_t e0 = guard.e0, e1 = guard.e1;
    // Time `steps` steps of tuning t from the snapshot, after one untimed step
    // and the capture of both step graphs; per-step milliseconds in *ms.
    auto trial = [&](const sph_tuning_t& t, int steps, float* ms) -> int {
        int r = sph_set_tuning(c, &t);
        if (r != SPH_OK) return r;
        if (hipMemcpyAsync(c->d.aos, c->d_backup, sizeof(sph_particle_t) * c->n, hipMemcpyDeviceToDevice,
                           c->stream) != hipSuccess)
            return hip_fail(c, hipErrorUnknown, "restore snapshot");
        invalidate_state(c);
        if ((r = sph_step(c, 1)) != SPH_OK) return r;
        if ((r = sph_step(c, 0)) != SPH_OK) return r;
        if (hipEventRecord(e0, c->stream) != hipSuccess) return hip_fail(c, hipErrorUnknown, "event");
        if ((r = sph_step(c, steps)) != SPH_OK) return r;
        float total = 0.f;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&total, e0, e1) != hipSuccess)
            return hip_fail(c, hipErrorUnknown, "autotune timing");
        *ms = total / (float)steps;
        return SPH_OK;
    };
    // Screen every setting with trial_steps steps, then re-time the four
    // fastest and the setting in force at the call with four times as many.
    // The fastest finalist replaces the setting in force only when it is more
    // than kTieMargin faster, so near-ties (timing noise between settings
    // whose step times differ by a percent) keep the caller's setting and
    // repeated calls agree.
    constexpr float kTieMargin = 0.02f;
    auto same = [](const sph_tuning_t& a, const sph_tuning_t& b) {
        return a.cells_per_h == b.cells_per_h && a.prune == b.prune && a.block == b.block &&
               a.blocks_per_xcd == b.blocks_per_xcd;
    };
    std::vector<std::pair<float, sph_tuning_t>> screened;
    const int32_t bpx_options[2] = {0, 64};
    for (int32_t sub = 1; sub <= 3 && rc == SPH_OK; ++sub)
        for (int32_t prune = 0; prune <= 1 && rc == SPH_OK; ++prune)
            for (int32_t block : {128, 256})
                for (int32_t bpx : bpx_options) {
                    if (rc != SPH_OK) break;
                    const sph_tuning_t t{sub, prune, block, bpx};
                    float ms = 0.f;
                    if ((rc = trial(t, trial_steps, &ms)) == SPH_OK) screened.emplace_back(ms, t);
                }
    std::stable_sort(screened.begin(), screened.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<sph_tuning_t> finalists;
    for (size_t k = 0; k < screened.size() && k < 4; ++k) finalists.push_back(screened[k].second);
    bool start_in = false;
    for (const auto& t : finalists) start_in = start_in || same(t, start_tuning);
    if (!start_in) finalists.push_back(start_tuning);
    sph_tuning_t win = start_tuning;
    float win_ms = 1e30f, start_ms = 1e30f;
    for (size_t k = 0; k < finalists.size() && rc == SPH_OK; ++k) {
        float ms = 0.f;
        if ((rc = trial(finalists[k], 4 * trial_steps, &ms)) != SPH_OK) break;
        if (same(finalists[k], start_tuning)) start_ms = ms;
        if (ms < win_ms) {
            win_ms = ms;
            win = finalists[k];
        }
    }
    if (rc == SPH_OK && !same(win, start_tuning) && win_ms > start_ms * (1.f - kTieMargin)) {
        win = start_tuning;
        win_ms = start_ms;
    }
    const int trial_rc = rc;
    const std::string trial_err = c->err;
    rc = sph_set_tuning(c, trial_rc == SPH_OK ? &win : &start_tuning);
    SPH_HIP(c, hipMemcpyAsync(c->d.aos, c->d_backup, sizeof(sph_particle_t) * c->n, hipMemcpyDeviceToDevice, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    invalidate_state(c);
    if (trial_rc != SPH_OK) {
        c->err = trial_err;
        return trial_rc;
    }
    if (rc != SPH_OK) return rc;
    if (best) *best = win;
    if (best_ms) *best_ms = (double)win_ms;
    return SPH_OK;
}

int sph_mlp_set(sph_ctx* c, const float* params, size_t n, const sph_mlp_norm_t* norm)
{
    if (!c) return SPH_ERR_INVALID;
    if (!params || n != (size_t)kMlpParams || !norm) return fail(c, SPH_ERR_INVALID, "need SPH_MLP_PARAMS weights and a norm");
    const float nv[4] = {norm->v_scale, norm->rho_scale, norm->pi_scale, norm->acc_scale};
    if (!finite_all(nv, 4) || !finite_all(params, (int)n)) return fail(c, SPH_ERR_INVALID, "non-finite weights or scales");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    if ((rc = ensure_mlp_base(c)) != SPH_OK) return rc;
    SPH_HIP(c, hipMemcpyAsync(c->mlp.params, params, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    SPH_HIP(c, hipMemsetAsync(c->mlp.adam_m, 0, sizeof(float) * kMlpParams, c->stream));
    SPH_HIP(c, hipMemsetAsync(c->mlp.adam_v, 0, sizeof(float) * kMlpParams, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    c->mlp.norm = MlpNorm{norm->v_scale, norm->rho_scale, norm->pi_scale, norm->acc_scale};
    c->mlp.pack_stale = true;
    c->mlp_loaded = true;
    c->adam_step = 0;
    return SPH_OK;
}

int sph_mlp_get(sph_ctx* c, float* params, size_t n, sph_mlp_norm_t* norm)
{
    if (!c) return SPH_ERR_INVALID;
    if (!c->mlp_loaded) return fail(c, SPH_ERR_STATE, "no network loaded");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    if (params) {
        if (n != (size_t)kMlpParams) return fail(c, SPH_ERR_INVALID, "need SPH_MLP_PARAMS entries");
        SPH_HIP(c, hipMemcpyAsync(params, c->mlp.params, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (norm) *norm = sph_mlp_norm_t{c->mlp.norm.v_scale, c->mlp.norm.rho_scale, c->mlp.norm.pi_scale, c->mlp.norm.acc_scale};
    return SPH_OK;
}

int sph_copy_state(sph_ctx* dst, const sph_ctx* src)
{
    if (!dst) return SPH_ERR_INVALID;
    if (!src || src->n != dst->n) return fail(dst, SPH_ERR_INVALID, "source must have the same particle count");
    int rc = set_device(dst);
    if (rc != SPH_OK) return rc;
    SPH_HIP(dst, hipStreamSynchronize(src->stream));
    SPH_HIP(dst, hipMemcpyAsync(dst->d.aos, src->d.aos, sizeof(sph_particle_t) * dst->n, hipMemcpyDeviceToDevice,
                                dst->stream));
    invalidate_state(dst);
    return SPH_OK;
}

int sph_mlp_train(sph_ctx* c, const sph_ctx* later, const sph_train_opts_t* o, float* grad_out, sph_train_result_t* res)
{
    if (!c) return SPH_ERR_INVALID;
    if (!o || !later || later->n != c->n || later->nscenes != c->nscenes)
        return fail(c, SPH_ERR_INVALID, "need options and a later state of the same particles");
    if (!c->mlp_loaded) return fail(c, SPH_ERR_STATE, "no network loaded (sph_mlp_set)");
    if (o->stride < 1 || o->first >= c->n) return fail(c, SPH_ERR_INVALID, "bad target range");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    // Training launches are never event-timed; the caller's setting is
    // restored on every way out.
    struct TimingOff {
        sph_ctx* c;
        bool was;
        ~TimingOff() { c->timing = was; }
    } timing_off{c, c->timing};
    c->timing = false;
    SPH_HIP(c, hipStreamSynchronize(later->stream));
    if ((rc = density_pass(c, true)) != SPH_OK) return rc;
    const uint32_t T = (c->n - o->first + o->stride - 1) / o->stride;
    if ((rc = ensure_targets(c, T)) != SPH_OK) return rc;
    MlpWork& w = c->mlp;
    SPH_HIP(c, launch_pairs(c->stream, c->d, c->nscenes, w, o->first, o->stride, T, false));
    std::vector<uint32_t> cnt(T);
    SPH_HIP(c, hipMemcpyAsync(cnt.data(), w.cnt, sizeof(uint32_t) * T, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<uint64_t> off(T + 1);
    off[0] = 0;
    for (uint32_t t = 0; t < T; ++t) off[t + 1] = off[t] + cnt[t];
    const uint64_t P = off[T];
    if ((rc = ensure_pairs(c, std::max<uint64_t>(P, 1))) != SPH_OK) return rc;
    const uint32_t R = 2048;
    const uint64_t nblocks = std::max<uint64_t>((P + R - 1) / R, 1);
    if (nblocks > w.partial_cap || !w.partial) {
        uint64_t cap = w.partial_cap;
        if ((rc = grow(c, &w.partial, cap, nblocks, kMlpParams, "gradient partials")) != SPH_OK) return rc;
        w.partial_cap = cap;
    }
    SPH_HIP(c, hipMemcpyAsync(w.off, off.data(), sizeof(uint64_t) * (T + 1), hipMemcpyHostToDevice, c->stream));
    SPH_HIP(c, launch_pairs(c->stream, c->d, c->nscenes, w, o->first, o->stride, T, true));
    SPH_HIP(c, launch_train_targets(c->stream, c->d, c->nscenes, w, later->d.aos, o->first, o->stride, T));
    std::vector<float> wt(T);
    SPH_HIP(c, hipMemcpyAsync(wt.data(), w.wt, sizeof(float) * T, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    double wsum = 0.0;
    for (float x : wt) wsum += x;
    const float inv_wsum = wsum > 0.0 ? (float)(1.0 / wsum) : 0.f;
    SPH_HIP(c, launch_mlp_forward(c->stream, w, P));
    SPH_HIP(c, launch_mlp_aggregate(c->stream, w, T, inv_wsum));
    SPH_HIP(c, launch_mlp_backward(c->stream, w, P, R, (uint32_t)nblocks));
    std::vector<double> lossp(T);
    SPH_HIP(c, hipMemcpyAsync(lossp.data(), w.lossp, sizeof(double) * T, hipMemcpyDeviceToHost, c->stream));
    if (grad_out)
        SPH_HIP(c, hipMemcpyAsync(grad_out, w.grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, c->stream));
    if (o->update) {
        c->adam_step += 1;
        const float bc1 = 1.f - std::pow(o->beta1, (float)c->adam_step);
        const float bc2 = 1.f - std::pow(o->beta2, (float)c->adam_step);
        SPH_HIP(c, launch_mlp_adam(c->stream, w, o->lr, o->beta1, o->beta2, o->eps, bc1, bc2));
        w.pack_stale = true;
    }
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    double L = 0.0;
    for (double x : lossp) L += x;
    if (res) {
        res->loss = wsum > 0.0 ? L / wsum : 0.0;
        res->pairs = P;
        res->targets = T;
        res->weighted = (uint32_t)wsum;
    }
    return SPH_OK;
}

int sph_mlp_predict(sph_ctx* c, int path, float* out)
{
    if (!c) return SPH_ERR_INVALID;
    if (!out) return fail(c, SPH_ERR_INVALID, "out is null");
    if (path != 0 && path != 1) return fail(c, SPH_ERR_INVALID, "path must be 0 (FP32) or 1 (MFMA)");
    if (!c->mlp_loaded) return fail(c, SPH_ERR_STATE, "no network loaded (sph_mlp_set)");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    if ((rc = density_pass(c, true)) != SPH_OK) return rc;
    if ((rc = ensure_targets(c, c->n)) != SPH_OK) return rc;
    if (!c->d_pred_aos && dalloc(&c->d_pred_aos, (size_t)c->n * 3u) != hipSuccess) {
        (void)hipGetLastError();
        c->d_pred_aos = nullptr;
        return fail(c, SPH_ERR_NOMEM, "cannot allocate prediction buffer");
    }
    if (path == 1 && c->mlp.pack_stale) {
        SPH_RUN(c, "mlp_pack", launch_mlp_pack(c->stream, c->mlp));
        c->mlp.pack_stale = false;
    }
    SPH_RUN(c, "mlp_messages", launch_coarse_mlp(c->stream, c->d, c->nscenes, c->mlp, path));
    SPH_HIP(c, launch_pred_to_aos(c->stream, c->d, c->mlp, c->d_pred_aos));
    SPH_HIP(c, hipMemcpyAsync(out, c->d_pred_aos, sizeof(float) * 3u * c->n, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    return SPH_OK;
}

namespace {

int mean_rho(sph_ctx* c, double* out)
{
    double tot = 0.0;
    for (int s = 0; s < c->nscenes; ++s) {
        sph_metrics_t m;
        int rc = sph_metrics(c, s, &m);
        if (rc != SPH_OK) return rc;
        tot += m.mean_rho * (double)c->count[s];
    }
    *out = tot / (double)c->n;
    return SPH_OK;
}

// One coarse step from the current AoS state.
int coarse_once(sph_ctx* c, bool use_network)
{
    int rc;
    if ((rc = density_pass(c, false)) != SPH_OK) return rc;
    if (use_network) {
        if (c->mlp.pack_stale) {
            SPH_RUN(c, "mlp_pack", launch_mlp_pack(c->stream, c->mlp));
            c->mlp.pack_stale = false;
        }
        SPH_RUN(c, "mlp_messages", launch_coarse_mlp(c->stream, c->d, c->nscenes, c->mlp, 1));
    } else {
        SPH_RUN(c, "force", launch_force(c->stream, c->d, ForceMode::Pass));
    }
    SPH_RUN(c, "coarse_integrate", launch_coarse_integrate(c->stream, c->d, c->mlp, !use_network));
    c->b_valid = true;
    c->a_valid = false;
    c->rho_valid = false;
    return SPH_OK;
}

int copy_aos(sph_ctx* c, sph_particle_t* dst, const sph_particle_t* src)
{
    SPH_HIP(c, hipMemcpyAsync(dst, src, sizeof(sph_particle_t) * c->n, hipMemcpyDeviceToDevice, c->stream));
    return SPH_OK;
}

}  // namespace

int sph_coarse_step(sph_ctx* c, int n, const sph_coarse_opts_t* o, sph_coarse_stats_t* st)
{
    if (!c) return SPH_ERR_INVALID;
    if (!o || n < 0 || o->judge_every < 0 || o->cooldown < 0 || !(o->drift_limit >= 0.f))
        return fail(c, SPH_ERR_INVALID, "bad coarse-step options");
    if (o->use_network && !c->mlp_loaded) return fail(c, SPH_ERR_STATE, "no network loaded (sph_mlp_set)");
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    if ((rc = ensure_targets(c, c->n)) != SPH_OK) return rc;  // per-slot predictions live in mlp.pred
    if (o->judge_every > 0 && (!c->d_judge_s || !c->d_judge_p)) {
        if (dalloc(&c->d_judge_s, c->n) != hipSuccess || dalloc(&c->d_judge_p, c->n) != hipSuccess) {
            (void)hipGetLastError();
            return fail(c, SPH_ERR_NOMEM, "cannot allocate judge buffers");
        }
    }
    sph_coarse_stats_t local{};
    sph_coarse_stats_t& s = st ? *st : local;
    s = sph_coarse_stats_t{};
    // After a rejected judge, `cooldown` coarse steps run as 4 plain steps
    // each and the step after them is judged again; a rejected re-judge starts
    // another cooldown, so while the network stays out of tolerance the
    // trajectory stays plain.
    int cooldown = 0;
    bool rejudge = false;
    for (int k = 0; k < n; ++k) {
        if (cooldown > 0) {
            // The cooldown's coarse steps (up to the call's end) as one call of
            // 4 plain steps each: the same trajectory in 16-step graphs.  At most
            // 2^28 coarse steps per call, so 4*run fits an int.
            const int run = std::min({cooldown, n - k, 1 << 28});
            cooldown -= run;
            k += run - 1;
            if ((rc = sph_step(c, 4 * run)) != SPH_OK) return rc;
            s.plain_steps += 4 * run;
            rejudge = cooldown == 0;
            continue;
        }
        const bool judge = rejudge || (o->judge_every > 0 && (k % o->judge_every) == 0);
        rejudge = false;
        if (!judge) {
            if ((rc = coarse_once(c, o->use_network != 0)) != SPH_OK) return rc;
            s.coarse_steps += 1;
            continue;
        }
        // Judge: 4 plain steps and one coarse step from the same state.
        double mp = 0.0, mc = 0.0;
        if ((rc = copy_aos(c, c->d_judge_s, c->d.aos)) != SPH_OK) return rc;
        invalidate_state(c);
        if ((rc = sph_step(c, 4)) != SPH_OK) return rc;
        if ((rc = density_pass(c, true)) != SPH_OK) return rc;
        if ((rc = mean_rho(c, &mp)) != SPH_OK) return rc;
        if ((rc = copy_aos(c, c->d_judge_p, c->d.aos)) != SPH_OK) return rc;
        if ((rc = copy_aos(c, c->d.aos, c->d_judge_s)) != SPH_OK) return rc;
        invalidate_state(c);
        if ((rc = coarse_once(c, o->use_network != 0)) != SPH_OK) return rc;
        if ((rc = density_pass(c, true)) != SPH_OK) return rc;
        if ((rc = mean_rho(c, &mc)) != SPH_OK) return rc;
        // Relative drift of the mean density against the plain result.
        const double drift = std::fabs(mc - mp) / std::fabs(mp);
        s.judged += 1;
        s.last_drift = drift;
        s.max_drift = std::max(s.max_drift, drift);
        if (!(drift <= (double)o->drift_limit)) {
            if ((rc = copy_aos(c, c->d.aos, c->d_judge_p)) != SPH_OK) return rc;
            invalidate_state(c);
            s.fallbacks += 1;
            s.plain_steps += 4;
            cooldown = o->cooldown;
        } else {
            s.coarse_steps += 1;
        }
    }
    if (o->use_network != 0 && n > 0) {
        unsigned long long pairs = 0;
        SPH_HIP(c, hipMemcpyAsync(&pairs, c->mlp.pairs, sizeof(pairs), hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(c, hipStreamSynchronize(c->stream));
        s.pairs = (int64_t)pairs;
    }
    return SPH_OK;
}

int sph_set_timing(sph_ctx* c, int enabled)
{
    if (!c) return SPH_ERR_INVALID;
    c->timing = enabled != 0;
    return SPH_OK;
}

int sph_kernel_times(sph_ctx* c, int max_entries, char (*names)[32], double* total_ms, int64_t* launches)
{
    if (!c) return SPH_ERR_INVALID;
    int rc = set_device(c);
    if (rc != SPH_OK) return rc;
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    for (auto& p : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.e0, p.e1) == hipSuccess) {
            auto& slot = c->times[p.name];
            slot.first += ms;
            slot.second += 1;
        } else {
            (void)hipGetLastError();
        }
        c->event_pool.push_back(p.e0);
        c->event_pool.push_back(p.e1);
    }
    c->pending.clear();
    int k = 0;
    for (auto& kv : c->times) {
        if (k >= max_entries) break;
        if (names) {
            std::snprintf(names[k], 32, "%s", kv.first.c_str());
        }
        if (total_ms) total_ms[k] = kv.second.first;
        if (launches) launches[k] = kv.second.second;
        ++k;
    }
    if (max_entries <= 0) return (int)c->times.size();
    if (!names && !total_ms && !launches) return (int)c->times.size();
    c->times.clear();
    return k;
}

}  // extern "C"
