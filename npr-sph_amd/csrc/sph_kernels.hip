// sph_kernels.hip - CDNA4 (gfx950) kernels of the SPH solver.
//
// One simulation step over cell-sorted state:
//   k_hash_count   cell key per particle + wave-aggregated per-cell counts
//   k_scan_*       exclusive scan of the counts -> first slot of every cell
//   k_scatter      arrival-order placement into cell ranges
//   k_rank_reorder stable order inside each cell (by AoS index) + gather of
//                  pos/vel into the sorted arrays the gathers read
//   k_density      rho_pres_comp.glsl:35-58 as a 27-cell gather
//   k_force        force_comp.glsl:35-66 as a 27-cell gather, optionally fused
//                  with integrate_comp.glsl:35-81 and the AoS write-back
//   k_integrate_aos integrate_comp.glsl:35-81 as a streaming pass over the AoS
// Scenes are independent; every launch covers all of them through the tile
// table (a tile is kTile consecutive particles of one scene).
#include <algorithm>

#include "sph_gather.h"
#include "sph_internal.h"

namespace nprsph {

namespace {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8), and each XCD has its own L2.  The gathers give each XCD one
// contiguous eighth of the cell-sorted slots (for_each_gather_tile), so its
// L2 holds that region's neighbours.
constexpr int kXcds = 8;

// Tiles per XCD.
__host__ __device__ __forceinline__ uint32_t xcd_per(uint32_t ntiles) { return (ntiles + kXcds - 1) / kXcds; }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t t = __shfl_up(v, (unsigned)d);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d);
    return v;
}

// Exclusive scan across a workgroup of NT threads; returns the workgroup total
// through `total`.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& total)
{
    constexpr int NW = NT / kWave;
    __shared__ uint32_t wtot[NW];
    const int lane = lane_id();
    const int w = threadIdx.x / kWave;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == kWave - 1) wtot[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const uint32_t t = wtot[k];
        before += (k < w) ? t : 0u;
        all += t;
    }
    __syncthreads();
    total = all;
    return before + inc - v;
}

template <int NT>
__device__ __forceinline__ uint32_t block_sum(uint32_t v)
{
    constexpr int NW = NT / kWave;
    __shared__ uint32_t wtot[NW];
    v = wave_sum(v);
    if (lane_id() == 0) wtot[threadIdx.x / kWave] = v;
    __syncthreads();
    uint32_t all = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) all += wtot[k];
    __syncthreads();
    return all;
}

// ------------------------------------------------------------ utilities ---

__global__ void k_set_scene(SceneDev* dst, SceneDev value) { *dst = value; }

__global__ __launch_bounds__(256) void k_fill_aos_defaults(sph_particle_t* __restrict__ aos, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float4* r = reinterpret_cast<float4*>(aos + i);
    r[0] = make_float4(0.f, 0.f, 0.f, 1.f);
    r[1] = make_float4(0.f, 0.f, 0.f, 0.f);
    r[2] = make_float4(0.f, 0.f, 0.f, 0.f);
    r[3] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t value)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ __launch_bounds__(256) void k_identity(uint32_t* __restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = i;
}

// Seeded jittered lattice (make_grid/init_particles, Main.cpp:488-539, with
// jitter).  Same float32 operation order as the CPU generator.
__global__ __launch_bounds__(256) void k_lattice(sph_particle_t* __restrict__ aos, uint32_t first, sph_lattice_t spec)
{
#pragma clang fp contract(off)
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= spec.count) return;
    const uint32_t nz = (uint32_t)spec.nz, ny = (uint32_t)spec.ny;
    const uint32_t idx[3] = {k / (nz * ny), (k / nz) % ny, k % nz};
    float p[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float base = spec.origin[a] + (float)idx[a] * spec.spacing;
        p[a] = base + lattice_jitter(spec.seed, 3ull * (uint64_t)k + (uint64_t)a, spec.jitter);
    }
    float4* r = reinterpret_cast<float4*>(aos + first + k);
    r[0] = make_float4(p[0], p[1], p[2], 1.f);
    r[1] = make_float4(0.f, 0.f, 0.f, 0.f);
    r[2] = make_float4(0.f, 0.f, 0.f, 0.f);
    r[3] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Copy AoS state into slot order B (pos/vel) and/or A (rho, p) through perm.
// The extras go to rho and the w lanes of A: PA.w = p, VA.w = 1/rho.
__global__ __launch_bounds__(kTile) void k_refresh(const Tile* __restrict__ tiles, const sph_particle_t* __restrict__ aos,
                                                   const uint32_t* __restrict__ perm, float4* __restrict__ P,
                                                   float4* __restrict__ V, float* __restrict__ rho,
                                                   float* __restrict__ PAw, float* __restrict__ VAw, int pos_vel,
                                                   int extras)
{
    const Tile t = tiles[blockIdx.x];
    const uint32_t s = t.begin + threadIdx.x;
    if (s >= t.end) return;
    const float4* r = reinterpret_cast<const float4*>(aos + perm[s]);
    if (pos_vel) {
        const float4 p = r[0], v = r[1];
        P[s] = make_float4(p.x, p.y, p.z, 0.f);
        V[s] = make_float4(v.x, v.y, v.z, 0.f);
    }
    if (extras) {
        const float4 e = r[3];
        const float inv = 1.0f / e.x;
        rho[s] = e.x;
        PAw[4 * (size_t)s + 3] = e.y;
        VAw[4 * (size_t)s + 3] = inv;
    }
}

// ------------------------------------------------------- neighbour grid ---

// Counts slot s (cell key k) into its cell and stores its arrival rank there.
// Called by the whole wave; lanes with active = false count nothing.  Lanes
// holding a run of equal keys (cell-sorted slots make runs the common case)
// share one atomic.
__device__ __forceinline__ void count_into_cell(uint32_t k, bool active, uint32_t s, uint32_t* __restrict__ slot,
                                                uint32_t* __restrict__ count, uint32_t* __restrict__ bsum)
{
    const int lane = lane_id();
    const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);  // bits 0..lane
    // Lane `lane` heads a run [lane, next) of equal values of v; returns next.
    auto run_end = [&](bool head) {
        const uint64_t after = __ballot(head) & ~upto;
        return after ? (__ffsll((long long)after) - 1) : kWave;
    };
    const uint32_t prev = __shfl_up(k, 1u);
    const bool head = !active || lane == 0 || prev != k;
    const int my_head = 63 - __clzll((long long)(__ballot(head) & upto));
    const int next = run_end(head);
    uint32_t base = 0;
    if (head && active) base = atomicAdd(&count[k], (uint32_t)(next - lane));
    base = __shfl(base, my_head);
    if (active) slot[s] = base + (uint32_t)(lane - my_head);
    // With bsum, the scan's block sums (kScanBlock cells each) are counted
    // alongside, so the build needs no reduction pass over the cells.  Worth it
    // only for cell-sorted slots, where a wave's keys fall in one or two blocks.
    if (!bsum) return;
    const uint32_t bk = k / (uint32_t)kScanBlock;
    const uint32_t bprev = __shfl_up(bk, 1u);
    const bool bhead = !active || lane == 0 || bprev != bk;
    const int bnext = run_end(bhead);
    if (bhead && active) atomicAdd(&bsum[bk], (uint32_t)(bnext - lane));
}

// Cell key of every slot of B, and its arrival rank inside the cell.  The
// fused force+integrate kernel does the same for the state it writes, and
// also counts the scan's block sums, so a step that follows one starts its
// build at the scan.  Here B may be in any order (after an upload), where
// per-particle block-sum atomics would contend, so k_scan_reduce sums them.
__global__ __launch_bounds__(kTile) void k_hash_count(const Tile* __restrict__ tiles, const SceneDev* __restrict__ scenes,
                                                      const float4* __restrict__ P, uint32_t* __restrict__ key,
                                                      uint32_t* __restrict__ slot, uint32_t* __restrict__ count)
{
    const Tile t = tiles[blockIdx.x];
    const SceneDev& sc = scenes[t.scene];
    const uint32_t s = t.begin + threadIdx.x;
    const bool active = s < t.end;
    uint32_t k = 0xffffffffu;
    if (active) {
        const float4 p = P[s];
        k = global_key(sc, p.x, p.y, p.z);
        key[s] = k;
    }
    count_into_cell(k, active, s, slot, count, nullptr);
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ count, uint32_t n,
                                                              uint32_t* __restrict__ bsum)
{
    const uint32_t base = blockIdx.x * (uint32_t)kScanBlock;
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t i = base + (uint32_t)(k * kScanThreads) + threadIdx.x;
        if (i < n) sum += count[i];
    }
    sum = block_sum<kScanThreads>(sum);
    if (threadIdx.x == 0) bsum[blockIdx.x] = sum;
}

__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* __restrict__ bsum, uint32_t nb)
{
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 1024u) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? bsum[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_excl_scan<1024>(v, total);
        if (i < nb) bsum[i] = carry + ex;
        carry += total;
    }
}

// start[c] = exclusive prefix of count; counts are cleared for the next build
// and start[n] receives the total.  Each thread owns kScanItems consecutive
// cells and moves them as 16-byte vectors where the whole vector is in range
// (count and start are 256-byte aligned, kScanItems is a multiple of 4).
// With fold, bsum holds the raw block sums and each block adds up those
// before it (grids of at most kScanFoldBlocks blocks, so at most that many
// reads per block); otherwise k_scan_top has already scanned them.
__global__ __launch_bounds__(kScanThreads) void k_scan_down(uint32_t* __restrict__ count, uint32_t n,
                                                            const uint32_t* __restrict__ bsum, uint32_t* __restrict__ start,
                                                            uint32_t fold)
{
    static_assert(kScanItems % 4 == 0, "vector width");
    const uint32_t base = blockIdx.x * (uint32_t)kScanBlock + threadIdx.x * (uint32_t)kScanItems;
    const bool whole = base + (uint32_t)kScanItems <= n;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k += 4) {
        const uint32_t i = base + (uint32_t)k;
        if (whole) {
            const uint4 q = *reinterpret_cast<const uint4*>(count + i);
            v[k] = q.x;
            v[k + 1] = q.y;
            v[k + 2] = q.z;
            v[k + 3] = q.w;
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) v[k + e] = i + (uint32_t)e < n ? count[i + (uint32_t)e] : 0u;
        }
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) sum += v[k];
    uint32_t prefix;
    if (fold) {
        uint32_t part = 0;
        for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kScanThreads) part += bsum[b];
        prefix = block_sum<kScanThreads>(part);
    } else {
        prefix = bsum[blockIdx.x];
    }
    uint32_t total;
    uint32_t run = prefix + block_excl_scan<kScanThreads>(sum, total);
    uint32_t out[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        out[k] = run;
        run += v[k];
    }
    if (whole) {
#pragma unroll
        for (int k = 0; k < kScanItems; k += 4) {
            *reinterpret_cast<uint4*>(start + base + k) = make_uint4(out[k], out[k + 1], out[k + 2], out[k + 3]);
            *reinterpret_cast<uint4*>(count + base + k) = make_uint4(0u, 0u, 0u, 0u);
        }
        if (base + (uint32_t)kScanItems == n) start[n] = run;
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const uint32_t i = base + (uint32_t)k;
            if (i < n) {
                start[i] = out[k];
                count[i] = 0u;
                if (i == n - 1u) start[n] = out[k] + v[k];
            }
        }
    }
}

// Also clears the scan's block sums, which the scan has consumed, for the
// next count.
__global__ __launch_bounds__(kTile) void k_scatter(const Tile* __restrict__ tiles, const uint32_t* __restrict__ key,
                                                   const uint32_t* __restrict__ slot, const uint32_t* __restrict__ start,
                                                   const uint32_t* __restrict__ perm, uint32_t* __restrict__ torig,
                                                   uint32_t* __restrict__ tsrc, uint32_t* __restrict__ tkey,
                                                   uint32_t* __restrict__ bsum, uint32_t nb)
{
    for (uint32_t i = blockIdx.x * (uint32_t)kTile + threadIdx.x; i < nb; i += gridDim.x * (uint32_t)kTile) bsum[i] = 0u;
    const Tile t = tiles[blockIdx.x];
    const uint32_t s = t.begin + threadIdx.x;
    if (s >= t.end) return;
    const uint32_t k = key[s];
    const uint32_t q = start[k] + slot[s];
    torig[q] = perm[s];
    tsrc[q] = s;
    tkey[q] = k;
}

// Final slot = cell start + number of cell members with a smaller AoS index,
// i.e. the order of a stable counting sort over AoS order, independent of the
// atomic arrival order.  Gathers pos/vel from B into A.  Members of the NaN
// cell interact with nothing, so they keep their arrival order (their slots
// are the only ones whose order is unspecified).
__global__ __launch_bounds__(kTile) void k_rank_reorder(const Tile* __restrict__ tiles, const SceneDev* __restrict__ scenes,
                                                        const uint32_t* __restrict__ start,
                                                        const uint32_t* __restrict__ torig, const uint32_t* __restrict__ tsrc,
                                                        const uint32_t* __restrict__ tkey, const float4* __restrict__ PB,
                                                        const float4* __restrict__ VB, uint32_t* __restrict__ perm,
                                                        float4* __restrict__ PA, float4* __restrict__ VA)
{
    const Tile t = tiles[blockIdx.x];
    const SceneDev& sc = scenes[t.scene];
    const uint32_t q = t.begin + threadIdx.x;
    if (q >= t.end) return;
    const uint32_t k = tkey[q];
    const uint32_t o = torig[q];
    uint32_t dst = q;
    if (k != sc.cell_base + sc.ncells) {
        const uint32_t b = start[k], e = start[k + 1];
        uint32_t rank = 0;
        for (uint32_t j = b; j < e; ++j) rank += (torig[j] < o) ? 1u : 0u;
        dst = b + rank;
    }
    const uint32_t src = tsrc[q];
    perm[dst] = o;
    PA[dst] = PB[src];
    VA[dst] = VB[src];
}

// ---------------------------------------------------------------- gathers ---

// Tiles of one gather launch: workgroup b takes the tiles (b % 8)*per + k for
// k = b/8, b/8 + G/8, ... (G = grid size, a multiple of 8), i.e. every XCD
// walks one contiguous eighth of the cell-sorted slots.  With G = 8*per each
// workgroup handles one tile; smaller G gives persistent workgroups.
template <class F>
__device__ __forceinline__ void for_each_gather_tile(uint32_t ntiles, F&& f)
{
    const uint32_t per = xcd_per(ntiles);
    const uint32_t stride = gridDim.x / kXcds;
    const uint32_t xcd = blockIdx.x % kXcds;
    for (uint32_t k = blockIdx.x / kXcds; k < per; k += stride) {
        const uint32_t ti = xcd * per + k;
        if (ti >= ntiles) break;
        f(ti);
    }
}

// rho_pres_comp.glsl:35-58.  rho_i = m*315/(64 pi h^9) * sum_{r<h} (h^2-r^2)^3
// over the stencil (self included, :46); p_i = max(k(rho_i-rho0), 0).
// Besides rho the pass stores p in P[s].w and 1/rho in V[s].w, so that the
// force pass (and the learned step) gets a neighbour's whole record in two
// 16-byte loads.  P[s] is rewritten whole, (x, y, z, p), from the position
// the lane loaded: a 16-byte store fills its slot instead of leaving a
// partly written one, and other waves reading that candidate concurrently
// see the same x, y, z bits before and after.  Candidates are read through
// the buffer resource of P; V's w lane is written through Vw (V as floats).
template <int SUB, bool PRUNE, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_density(const Tile* __restrict__ tiles, const SceneDev* __restrict__ scenes,
                                                   const uint32_t* __restrict__ start, float4* P,
                                                   float* __restrict__ rho_out, float* Vw,
                                                   const uint32_t* __restrict__ perm,
                                                   sph_particle_t* __restrict__ aos, uint32_t np, uint32_t ntiles)
{
    __shared__ uint32_t lds[BLOCK / 64][gather_lds_words<SUB>()];
    const __amdgpu_buffer_rsrc_t rP = f4_rsrc(P, np);
    for_each_gather_tile(ntiles, [&](uint32_t ti) {
        const Tile t = tiles[ti];
        const SceneDev& sc = scenes[t.scene];
        const uint32_t s = t.begin + threadIdx.x;
        const bool active = s < t.end;
        const float4 xi = P[active ? s : t.begin];
        const float h2 = sc.h2;
        float acc = 0.f;
        int cx = 0, cy = 0, cz = 0;
        float fx = 0.f, fy = 0.f, fz = 0.f;
        // Term (h^2-r^2)^3 of one candidate, h^2-r^2 formed as h^2-dz^2-dy^2-dx^2
        // by fused steps; r >= h and NaN/Inf distances add nothing.
        auto term = [&](const float4 xj) {
            const float ddx = xi.x - xj.x, ddy = xi.y - xj.y, ddz = xi.z - xj.z;
            const float w = fmaxf(fmaf(-ddx, ddx, fmaf(-ddy, ddy, fmaf(-ddz, ddz, h2))), 0.f);
            return w * w * w;
        };
        const bool valid = active && cell_coords(sc, xi.x, xi.y, xi.z, fx, fy, fz, cx, cy, cz);
        gather_rows<SUB, PRUNE, true>(sc, start, P, rP, lds[threadIdx.x / 64], valid, fx, fy, fz, cx, cy, cz,
                                [&](auto q, uint32_t m, uint32_t, bool) {
                                    for (; m >= 4u; m -= 4u, q += 4) {
                                        const float4 x0 = q[0], x1 = q[1], x2 = q[2], x3 = q[3];
                                        acc += term(x0);
                                        acc += term(x1);
                                        acc += term(x2);
                                        acc += term(x3);
                                    }
                                    for (; m > 0u; --m, ++q) acc += term(*q);
                                });
        if (!active) return;
        const float rho = acc * sc.poly6_m;
        const float p = fmaxf(sc.gas_const * (rho - sc.rho0), 0.f);
        const float inv = 1.0f / rho;
        rho_out[s] = rho;
        P[s] = make_float4(xi.x, xi.y, xi.z, p);
        Vw[4 * (size_t)s + 3] = inv;
        if (aos) {
            float* e = aos[perm[s]].extras;
            e[0] = rho;
            e[1] = p;
        }
    });
}

// Hits found while scanning candidates are queued per lane in LDS (a ring of
// kQueue entries) and then processed by all lanes in lock step, so the
// pair-force arithmetic runs once per neighbour instead of once per candidate
// of the slowest lane.  Every candidate's index is written to the lane's next
// queue entry and the entry is kept (tail advanced) only for a hit, so the
// scan has no branches.  A drain stops once fewer than kDrainLanes lanes have
// entries left, provided no lane is within 8 entries of full: the lanes with
// long queues keep their entries for a later drain instead of running alone.
constexpr int kQueue = 16;
constexpr int kDrainLanes = 32;
static_assert((kQueue & (kQueue - 1)) == 0 && kQueue >= 16, "queue ring size");

// One neighbour's pressure and viscosity terms for particle (xi, vi, pi),
// accumulated into a (pressure) and b (viscosity).  The hardware reciprocal
// square root of r^2 = 0 (a coincident pair) is +inf, so r = 0*inf is NaN, as
// normalize(0) in the reference.
struct PairSums {
    float4 xi, vi;
    float h;
    float ax = 0.f, ay = 0.f, az = 0.f;
    float bx = 0.f, by = 0.f, bz = 0.f;
    __device__ __forceinline__ void add(const float4 xj, const float4 vj)
    {
        const float ddx = xi.x - xj.x, ddy = xi.y - xj.y, ddz = xi.z - xj.z;
        const float r2 = fmaf(ddx, ddx, fmaf(ddy, ddy, ddz * ddz));
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        const float r = r2 * inv_r;
        const float q = h - r;
        const float cp = (xi.w + xj.w) * vj.w * (q * q) * inv_r;
        ax = fmaf(cp, ddx, ax);
        ay = fmaf(cp, ddy, ay);
        az = fmaf(cp, ddz, az);
        const float cv = vj.w * q;
        bx = fmaf(cv, vj.x - vi.x, bx);
        by = fmaf(cv, vj.y - vi.y, by);
        bz = fmaf(cv, vj.z - vi.z, bz);
    }
};

// F = pres_m*a + visc_m*b + rho_i*G (force_comp.glsl:60-66), then either the
// force into the AoS record (the force pass) or, with FUSE,
// integrate_comp.glsl:35-81 on it: the next step's B state, its count into
// the next grid's cells (as k_hash_count would from B) and, unless aos is
// null, the AoS record (pos, vel, force, extras.xy).  Called by every lane
// of the wave (the cell count takes the whole wave).
template <bool FUSE>
__device__ __forceinline__ void force_epilogue(const SceneDev& sc, uint32_t s, bool active, const PairSums& ps,
                                               float rhoi, const uint32_t* __restrict__ perm,
                                               sph_particle_t* __restrict__ aos, float4* __restrict__ PN,
                                               float4* __restrict__ VN, uint32_t* __restrict__ key,
                                               uint32_t* __restrict__ slot, uint32_t* __restrict__ count,
                                               uint32_t* __restrict__ bsum)
{
    float f[3];
    f[0] = fmaf(sc.pres_m, ps.ax, fmaf(sc.visc_m, ps.bx, rhoi * sc.grav[0]));
    f[1] = fmaf(sc.pres_m, ps.ay, fmaf(sc.visc_m, ps.by, rhoi * sc.grav[1]));
    f[2] = fmaf(sc.pres_m, ps.az, fmaf(sc.visc_m, ps.bz, rhoi * sc.grav[2]));
    if constexpr (FUSE) {
        const float x[3] = {ps.xi.x, ps.xi.y, ps.xi.z}, v[3] = {ps.vi.x, ps.vi.y, ps.vi.z};
        float nx[3], nv[3];
        integrate_one(sc, x, v, f, rhoi, nx, nv);
        uint32_t k = 0xffffffffu;
        if (active) {
            PN[s] = make_float4(nx[0], nx[1], nx[2], 0.f);
            VN[s] = make_float4(nv[0], nv[1], nv[2], 0.f);
            k = global_key(sc, nx[0], nx[1], nx[2]);
            key[s] = k;
        }
        count_into_cell(k, active, s, slot, count, bsum);  // the whole wave takes part
        if (!active || !aos) return;  // no record, or one a later step of the same call overwrites
        float* r = reinterpret_cast<float*>(aos + perm[s]);
        r[0] = nx[0];
        r[1] = nx[1];
        r[2] = nx[2];
        r[4] = nv[0];
        r[5] = nv[1];
        r[6] = nv[2];
        r[8] = f[0];
        r[9] = f[1];
        r[10] = f[2];
        r[12] = rhoi;
        r[13] = ps.xi.w;
    } else {
        if (!active) return;
        float* r = reinterpret_cast<float*>(aos + perm[s]);
        r[8] = f[0];
        r[9] = f[1];
        r[10] = f[2];
    }
}

// force_comp.glsl:35-66 over the stencil, j != i (:50-53), r < h (:57):
//   pressure  += m(p_i+p_j)/(2 rho_j) * 45/(pi h^6) * (h-r)^2 * delta/r
//   viscosity += m(v_j-v_i)/rho_j * 45/(pi h^6) * (h-r), then * visc
//   F = pressure + viscosity + rho_i * G
// Reads P = (x, y, z, p) and V = (vx, vy, vz, 1/rho) as left by the density
// pass.  Slot `sentinel` (one past the last particle) holds a far-away
// particle with p = 0 and 1/rho = 0: its pair terms are exact zeros, and it
// fills the second entry of a queue pair that a lane does not have.  P and V
// hold sentinel + 1 elements.
// FUSE additionally runs integrate_comp.glsl:35-81 on the new force, writes the
// next step's B state, counts it into the next grid's cells (as k_hash_count
// would from B) and, unless aos is null, writes the AoS record (pos, vel,
// force, extras.xy).  No kernel of a step reads the AoS, so only the last step
// of a multi-step call needs to write it.
// Occupancy: at 1 or 2 cells per h the kernel fits 96 VGPRs (5 waves per SIMD
// instead of 4) with a few spills outside the neighbour loops, which is ~8%
// faster; at 3 cells per h the larger stencil would spill inside them.
template <bool FUSE, int SUB, bool PRUNE, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(SUB < 3 ? 5 : 4))) void k_force(const Tile* __restrict__ tiles, const SceneDev* __restrict__ scenes,
                                                 const uint32_t* __restrict__ start, const float4* __restrict__ P,
                                                 const float4* __restrict__ V, const float* __restrict__ rho,
                                                 const uint32_t* __restrict__ perm, sph_particle_t* __restrict__ aos,
                                                 float4* __restrict__ PN, float4* __restrict__ VN,
                                                 uint32_t* __restrict__ key, uint32_t* __restrict__ slot,
                                                 uint32_t* __restrict__ count, uint32_t* __restrict__ bsum,
                                                 uint32_t sentinel, uint32_t ntiles)
{
    __shared__ uint32_t queue[kQueue][BLOCK];
    __shared__ uint32_t lds[BLOCK / 64][gather_lds_words<SUB>()];
    const __amdgpu_buffer_rsrc_t rP = f4_rsrc(P, sentinel + 1u), rV = f4_rsrc(V, sentinel + 1u);
    for_each_gather_tile(ntiles, [&](uint32_t ti) {
        const Tile t = tiles[ti];
        const SceneDev& sc = scenes[t.scene];
        const uint32_t s = t.begin + threadIdx.x;
        const bool active = s < t.end;
        const uint32_t si = active ? s : t.begin;
        const float4 xi = P[si];
        const float4 vi = V[si];
        const float rhoi = rho[si];
        const float h2 = sc.h2;
        PairSums ps{xi, vi, sc.h};
        // This lane's queue entries are [head, tail) mod kQueue, both counted
        // in bytes of a queue row (kRow per entry), so an entry's LDS address
        // is one and-or with the lane's column; head is an even entry.
        static_assert((BLOCK & (BLOCK - 1)) == 0, "the ring wraps with the 32-bit byte counters");
        constexpr uint32_t kRow = BLOCK * 4u, kWrap = (kQueue - 1) * kRow;
        char* const qbase = reinterpret_cast<char*>(&queue[0][0]);
        const uint32_t col = threadIdx.x * 4u;
        uint32_t head = 0, tail = 0;
        // Processes queue entries two per round (the loads of both are issued
        // before either is used), oldest first.  With all = false it returns
        // with every lane at most kQueue - 8 entries full.  head stays even, so
        // a round's two entries are adjacent slots (one ds_read2); a lane that
        // had one entry moves its tail past the unused slot.
        auto flush = [&](bool all) {
            for (;;) {
                const uint32_t n = tail - head;  // bytes
                if (all ? !__any(n > 0u)
                        : (!__any(n > (kQueue - 8) * kRow) && __popcll(__ballot(n > 0u)) < kDrainLanes))
                    break;
                if (n > 0u) {
                    const uint32_t* e = reinterpret_cast<const uint32_t*>(qbase + ((head & kWrap) | col));
                    const uint32_t j0 = e[0], e1 = e[BLOCK];
                    const uint32_t j1 = n > kRow ? e1 : sentinel;
                    const float4 x0 = f4_load(rP, j0), v0 = f4_load(rV, j0);
                    const float4 x1 = f4_load(rP, j1), v1 = f4_load(rV, j1);
                    ps.add(x0, v0);
                    ps.add(x1, v1);
                    head += 2u * kRow;
                    tail = n > kRow ? tail : head;
                }
            }
        };
        // Needs a free entry (fewer than kQueue queued).  j is never this lane's
        // own slot: the scan below leaves it out of the candidate ranges.
        auto consider = [&](const float4 xj, uint32_t j) {
            const float ddx = xi.x - xj.x, ddy = xi.y - xj.y, ddz = xi.z - xj.z;
            const float r2 = fmaf(ddx, ddx, fmaf(ddy, ddy, ddz * ddz));
            *reinterpret_cast<uint32_t*>(qbase + ((tail & kWrap) | col)) = j;
            tail += r2 < h2 ? kRow : 0u;
        };
        int cx = 0, cy = 0, cz = 0;
        float fx = 0.f, fy = 0.f, fz = 0.f;
        const bool valid = active && cell_coords(sc, xi.x, xi.y, xi.z, fx, fy, fz, cx, cy, cz);
        gather_rows<SUB, PRUNE, true>(sc, start, P, rP, lds[threadIdx.x / 64], valid, fx, fy, fz, cx, cy, cz,
                                [&](auto q, uint32_t m, uint32_t j, bool self_row) {
                                    // Candidates [0, m) as [0, m0) and [m1, m): in the row
                                    // holding slot s, m0 = s - j and m1 = m0 + 1.
                                    uint32_t m0 = m, m1 = m;
                                    if (self_row && s - j < m) {
                                        m0 = s - j;
                                        m1 = m0 + 1u;
                                    }
#pragma nounroll
                                    for (int part = 0; part < (self_row ? 2 : 1); ++part) {
                                        const uint32_t first = part ? m1 : 0u;
                                        uint32_t n = part ? m - m1 : m0;
                                        auto c = q + first;
                                        uint32_t jc = j + first;
                                        for (; n >= 4u; n -= 4u, c += 4, jc += 4u) {
                                            const float4 x0 = c[0], x1 = c[1], x2 = c[2], x3 = c[3];
                                            if (__any(tail - head > (kQueue - 4) * kRow)) flush(false);
                                            consider(x0, jc);
                                            consider(x1, jc + 1u);
                                            consider(x2, jc + 2u);
                                            consider(x3, jc + 3u);
                                        }
                                        for (; n > 0u; --n, ++c, ++jc) {
                                            const float4 xj = *c;
                                            if (__any(tail - head == kQueue * kRow)) flush(false);
                                            consider(xj, jc);
                                        }
                                    }
                                });
        flush(true);
        force_epilogue<FUSE>(sc, s, active, ps, rhoi, perm, aos, PN, VN, key, slot, count, bsum);
    });
}

// Value of lane K of this lane's quad (DPP quad permute); every lane of the
// quad must be active.
template <int K>
__device__ __forceinline__ float quad_lane(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), K * 0x55, 0xF, 0xF, false));
}

// integrate_comp.glsl:35-81 as a streaming pass over the AoS buffer.  Lane q of
// a quad reads float4 q of one 64-byte record (pos, vel, force, extras), so a
// wave's loads are one contiguous 1 KB run.  The quad shares the record
// through DPP, every lane applies integrate_one, and lanes 0 and 1 write pos
// and vel back with their w lanes unchanged.
__global__ __launch_bounds__(kTile) void k_integrate_aos(const Tile* __restrict__ tiles, const SceneDev* __restrict__ scenes,
                                                         sph_particle_t* __restrict__ aos)
{
    static_assert(kTile % 4 == 0, "four lanes per record");
    const Tile t = tiles[blockIdx.x];
    const SceneDev& sc = scenes[t.scene];
    float4* r4 = reinterpret_cast<float4*>(aos);
    const uint32_t q = threadIdx.x & 3u;
    // All four rounds' loads are issued before the first store; quads past
    // the tile's end read its first record instead and store nothing.
    float4 rec[4];
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k) {
        const uint32_t i = t.begin + k * (uint32_t)(kTile / 4) + threadIdx.x / 4u;
        rec[k] = r4[4 * (size_t)(i < t.end ? i : t.begin) + q];
    }
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k) {
        const uint32_t i = t.begin + k * (uint32_t)(kTile / 4) + threadIdx.x / 4u;
        const bool live = i < t.end;  // uniform across the quad
        const float4 v = rec[k];
        const float x[3] = {quad_lane<0>(v.x), quad_lane<0>(v.y), quad_lane<0>(v.z)};
        const float vel[3] = {quad_lane<1>(v.x), quad_lane<1>(v.y), quad_lane<1>(v.z)};
        const float f[3] = {quad_lane<2>(v.x), quad_lane<2>(v.y), quad_lane<2>(v.z)};
        const float rho = quad_lane<3>(v.x);
        float nx[3], nv[3];
        integrate_one(sc, x, vel, f, rho, nx, nv);
        if (live && q < 2u)
            r4[4 * (size_t)i + q] = q == 0u ? make_float4(nx[0], nx[1], nx[2], v.w) : make_float4(nv[0], nv[1], nv[2], v.w);
    }
}

// ------------------------------------------------------------- reductions ---

// Per-tile partial sums over the AoS: rho, kinetic energy, position, max
// speed, count of non-finite pos/vel (kinetic energy, position and speed over
// finite particles only).  Eight doubles per tile.
// Tile first_tile + b writes its partials to partials[8b .. 8b+7].
__global__ __launch_bounds__(kTile) void k_metrics(const Tile* __restrict__ tiles, const SceneDev* __restrict__ scenes,
                                                   const sph_particle_t* __restrict__ aos, double* __restrict__ partials,
                                                   int first_tile)
{
    const int ti = first_tile + (int)blockIdx.x;
    const Tile t = tiles[ti];
    const SceneDev& sc = scenes[t.scene];
    const uint32_t i = t.begin + threadIdx.x;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (i < t.end) {
        const float4* r = reinterpret_cast<const float4*>(aos + i);
        const float4 p = r[0], vel = r[1], e = r[3];
        const double s2 = (double)vel.x * vel.x + (double)vel.y * vel.y + (double)vel.z * vel.z;
        const bool finite = isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(vel.x) && isfinite(vel.y) &&
                            isfinite(vel.z);
        v[0] = e.x;
        v[1] = finite ? 0.5 * (double)sc.mass * s2 : 0.0;
        v[2] = finite ? p.x : 0.0;
        v[3] = finite ? p.y : 0.0;
        v[4] = finite ? p.z : 0.0;
        v[5] = finite ? sqrt(s2) : 0.0;
        v[6] = finite ? 0.0 : 1.0;
    }
    __shared__ double red[8][kTile / 64];
    const int lane = lane_id(), w = threadIdx.x / 64;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        double x = v[k];
        for (int d = 32; d > 0; d >>= 1) {
            const double o = __shfl_xor(x, d);
            x = (k == 5) ? fmax(x, o) : x + o;
        }
        if (lane == 0) red[k][w] = x;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        const int k = threadIdx.x;
        double x = red[k][0];
        for (int q = 1; q < kTile / 64; ++q) x = (k == 5) ? fmax(x, red[k][q]) : x + red[k][q];
        partials[(size_t)blockIdx.x * 8 + k] = x;
    }
}

// out[k] = the ntiles per-tile partials of metric k combined in tile order
// (sum, or max for k = 5), one lane per metric: the same association as the
// host-side combination in sph_metrics, so both give the same bits.
__global__ __launch_bounds__(64) void k_metrics_fold(const double* __restrict__ partials, int ntiles,
                                                     double* __restrict__ out)
{
    const int k = threadIdx.x;
    if (k >= 7) return;
    double acc = 0.0;
    for (int t = 0; t < ntiles; ++t) {
        const double x = partials[(size_t)t * 8 + k];
        acc = (k == 5) ? (x > acc ? x : acc) : acc + x;
    }
    out[k] = acc;
}

inline unsigned blocks_for(uint32_t n) { return (n + 255u) / 256u; }

}  // namespace

// --------------------------------------------------------------- launchers ---

hipError_t launch_set_scene(hipStream_t s, SceneDev* dst, const SceneDev& value)
{
    hipLaunchKernelGGL(k_set_scene, dim3(1), dim3(1), 0, s, dst, value);
    return hipGetLastError();
}

hipError_t launch_fill_aos_defaults(hipStream_t s, sph_particle_t* aos, uint32_t n)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_aos_defaults, dim3(blocks_for(n)), dim3(256), 0, s, aos, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t n, uint32_t value)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n)), dim3(256), 0, s, p, n, value);
    return hipGetLastError();
}

hipError_t launch_identity(hipStream_t s, uint32_t* p, uint32_t n)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_identity, dim3(blocks_for(n)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_lattice(hipStream_t s, sph_particle_t* aos, uint32_t first, const sph_lattice_t& spec)
{
    if (spec.count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lattice, dim3(blocks_for(spec.count)), dim3(256), 0, s, aos, first, spec);
    return hipGetLastError();
}

hipError_t launch_refresh(hipStream_t s, const DeviceState& d, bool pos_vel, bool extras)
{
    hipLaunchKernelGGL(k_refresh, dim3(d.ntiles), dim3(kTile), 0, s, d.tiles, d.aos, d.perm, d.PB, d.VB, d.rho,
                       reinterpret_cast<float*>(d.PA), reinterpret_cast<float*>(d.VA), pos_vel ? 1 : 0,
                       extras ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_hash_count(hipStream_t s, const DeviceState& d)
{
    hipLaunchKernelGGL(k_hash_count, dim3(d.ntiles), dim3(kTile), 0, s, d.tiles, d.scenes, d.PB, d.key, d.slot, d.count);
    return hipGetLastError();
}

hipError_t launch_scan_reduce(hipStream_t s, const DeviceState& d)
{
    hipLaunchKernelGGL(k_scan_reduce, dim3(scan_blocks(d.total_cells)), dim3(kScanThreads), 0, s, d.count,
                       d.total_cells, d.bsum);
    return hipGetLastError();
}

hipError_t launch_scan_top(hipStream_t s, const DeviceState& d)
{
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, d.bsum, (uint32_t)scan_blocks(d.total_cells));
    return hipGetLastError();
}

hipError_t launch_scan_down(hipStream_t s, const DeviceState& d)
{
    hipLaunchKernelGGL(k_scan_down, dim3(scan_blocks(d.total_cells)), dim3(kScanThreads), 0, s, d.count, d.total_cells,
                       d.bsum, d.start, scan_folds(d.total_cells) ? 1u : 0u);
    return hipGetLastError();
}

hipError_t launch_scatter(hipStream_t s, const DeviceState& d)
{
    hipLaunchKernelGGL(k_scatter, dim3(d.ntiles), dim3(kTile), 0, s, d.tiles, d.key, d.slot, d.start, d.perm, d.torig,
                       d.tsrc, d.tkey, d.bsum, (uint32_t)scan_blocks(d.total_cells));
    return hipGetLastError();
}

hipError_t launch_rank_reorder(hipStream_t s, const DeviceState& d)
{
    hipLaunchKernelGGL(k_rank_reorder, dim3(d.ntiles), dim3(kTile), 0, s, d.tiles, d.scenes, d.start, d.torig, d.tsrc,
                       d.tkey, d.PB, d.VB, d.perm, d.PA, d.VA);
    return hipGetLastError();
}

namespace {

inline dim3 gather_grid(const DeviceState& d)
{
    const uint32_t per = xcd_per((uint32_t)d.ngtiles);
    const uint32_t bpx = d.tuning.blocks_per_xcd > 0 ? std::min<uint32_t>(per, (uint32_t)d.tuning.blocks_per_xcd) : per;
    return dim3(kXcds * std::max<uint32_t>(bpx, 1u));
}

template <int SUB, bool PRUNE, int BLOCK>
hipError_t density_variant(hipStream_t s, const DeviceState& d, int write_aos)
{
    hipLaunchKernelGGL((k_density<SUB, PRUNE, BLOCK>), gather_grid(d), dim3(BLOCK), 0, s, d.gtiles, d.scenes, d.start,
                       d.PA, d.rho, reinterpret_cast<float*>(d.VA), d.perm, write_aos ? d.aos : nullptr, d.n,
                       (uint32_t)d.ngtiles);
    return hipGetLastError();
}

// mode: ForcePass, or FusedStep with or without the AoS write.
template <int SUB, bool PRUNE, int BLOCK>
hipError_t force_variant(hipStream_t s, const DeviceState& d, int mode)
{
    if (mode != (int)ForceMode::Pass)
        hipLaunchKernelGGL((k_force<true, SUB, PRUNE, BLOCK>), gather_grid(d), dim3(BLOCK), 0, s, d.gtiles, d.scenes,
                           d.start, d.PA, d.VA, d.rho, d.perm, mode == (int)ForceMode::FusedAos ? d.aos : nullptr,
                           d.PB, d.VB, d.key, d.slot, d.count, d.bsum, d.n, (uint32_t)d.ngtiles);
    else
        hipLaunchKernelGGL((k_force<false, SUB, PRUNE, BLOCK>), gather_grid(d), dim3(BLOCK), 0, s, d.gtiles, d.scenes,
                           d.start, d.PA, d.VA, d.rho, d.perm, d.aos, d.PB, d.VB, nullptr, nullptr, nullptr, nullptr,
                           d.n, (uint32_t)d.ngtiles);
    return hipGetLastError();
}

using GatherFn = hipError_t (*)(hipStream_t, const DeviceState&, int);

#define NPRSPH_VARIANTS(fn)                                                                    \
    {{{fn<1, false, 128>, fn<1, false, 256>}, {fn<1, true, 128>, fn<1, true, 256>}},           \
     {{fn<2, false, 128>, fn<2, false, 256>}, {fn<2, true, 128>, fn<2, true, 256>}},           \
     {{fn<3, false, 128>, fn<3, false, 256>}, {fn<3, true, 128>, fn<3, true, 256>}}}

const GatherFn kDensity[3][2][2] = NPRSPH_VARIANTS(density_variant);
const GatherFn kForce[3][2][2] = NPRSPH_VARIANTS(force_variant);

GatherFn pick(const GatherFn (&table)[3][2][2], const sph_tuning_t& t)
{
    return table[t.cells_per_h - 1][t.prune ? 1 : 0][t.block == 128 ? 0 : 1];
}

}  // namespace

bool tuning_supported(const sph_tuning_t& t)
{
    return t.cells_per_h >= 1 && t.cells_per_h <= 3 && (t.prune == 0 || t.prune == 1) &&
           (t.block == 128 || t.block == 256) && t.blocks_per_xcd >= 0;
}

hipError_t launch_density(hipStream_t s, const DeviceState& d, bool write_aos)
{
    return pick(kDensity, d.tuning)(s, d, write_aos ? 1 : 0);
}

hipError_t launch_force(hipStream_t s, const DeviceState& d, ForceMode mode)
{
    return pick(kForce, d.tuning)(s, d, (int)mode);
}

hipError_t launch_integrate_aos(hipStream_t s, const DeviceState& d)
{
    hipLaunchKernelGGL(k_integrate_aos, dim3(d.ntiles), dim3(kTile), 0, s, d.tiles, d.scenes, d.aos);
    return hipGetLastError();
}

hipError_t launch_metrics(hipStream_t s, const DeviceState& d, int first_tile, int ntiles, double* partials)
{
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_metrics, dim3(ntiles), dim3(kTile), 0, s, d.tiles, d.scenes, d.aos, partials, first_tile);
    return hipGetLastError();
}

hipError_t launch_metrics_fold(hipStream_t s, const double* partials, int ntiles, double* out)
{
    hipLaunchKernelGGL(k_metrics_fold, dim3(1), dim3(64), 0, s, partials, ntiles, out);
    return hipGetLastError();
}

}  // namespace nprsph
