// sph_gather.h - neighbour enumeration shared by the gather kernels
// (density, force) and the learned coarse step's pair kernels.
#pragma once

#include "sph_device.h"

namespace nprsph {

// Visits the stencil rows of a particle at fractional cell coordinates
// (fx,fy,fz) in cell (cx,cy,cz): rows (y,z) within SUB cells, each a slot
// range [b,e) covering the x cells within SUBX = kXSplit*SUB.  With PRUNE,
// rows whose y/z gap to the particle is >= SUB cells are skipped and the x
// extent is cut to floor(fx -+ kXSplit*w), w = sqrt(SUB^2 - gap^2) in y/z
// cell edges (kXSplit*w in the kXSplit times narrower x cells; the factor is
// exact).  SUB y/z cells and SUBX x cells exceed h by the factor
// 1+2^-10, which covers the float32 rounding of the gaps and of the hardware
// square root, so pruning only drops pairs with r >= h.  Border cells also
// hold clamped particles beyond the box: gaps measured from the unclamped
// coordinates stay lower bounds for them, and the x window always keeps a
// border cell that its (unbounded) extent reaches.
// The row bounds of one z-plane are fetched together (one memory latency per
// plane) into the caller's LDS table `rows` (2*(2*SUB+1) words per thread,
// stride BLOCK).
template <int SUB, bool PRUNE, int BLOCK, class Body>
__device__ __forceinline__ void for_each_row(const SceneDev& sc, const uint32_t* __restrict__ start, float fx, float fy,
                                             float fz, int cx, int cy, int cz, uint32_t* rows, Body&& body)
{
    constexpr int kRows = 2 * SUB + 1;
    constexpr int SUBX = kXSplit * SUB;
    const int dx = sc.dims[0], dy = sc.dims[1], dz = sc.dims[2];
    const int xa = cx > SUBX ? cx - SUBX : 0;
    const int xb = cx < dx - SUBX ? cx + SUBX : dx - 1;
    const int z0 = cz > SUB ? cz - SUB : 0, z1 = cz < dz - SUB ? cz + SUB : dz - 1;
    constexpr float kReach2 = (float)(SUB * SUB);
    for (int z = z0; z <= z1; ++z) {
        const float gz = z > cz ? (float)z - fz : (z < cz ? fz - (float)(z + 1) : 0.f);
#pragma unroll
        for (int k = 0; k < kRows; ++k) {
            const int y = cy - SUB + k;
            uint32_t b = 0, e = 0;
            if (y >= 0 && y < dy) {
                int x0 = xa, x1 = xb;
                bool keep = true;
                if constexpr (PRUNE) {
                    const float gy = y > cy ? (float)y - fy : (y < cy ? fy - (float)(y + 1) : 0.f);
                    const float rem = kReach2 - fmaf(gy, gy, gz * gz);
                    keep = rem > 0.f;
                    const float w = (float)kXSplit * __builtin_amdgcn_sqrtf(fmaxf(rem, 0.f));
                    // Cells 0 and dx-1 extend to -inf/+inf, so the window ends
                    // are clamped to them before meeting the stencil bounds.
                    x0 = (int)fmaxf(fminf(floorf(fx - w), (float)(dx - 1)), (float)xa);
                    x1 = (int)fminf(fmaxf(floorf(fx + w), 0.f), (float)xb);
                }
                if (keep) {
                    const uint32_t row = row_key(sc, y, z);
                    b = start[row + (uint32_t)x0];
                    e = start[row + (uint32_t)(x1 + 1)];
                }
            }
            rows[(2 * k) * BLOCK] = b;
            rows[(2 * k + 1) * BLOCK] = e;
        }
        for (int k = 0; k < kRows; ++k) body(rows[(2 * k) * BLOCK], rows[(2 * k + 1) * BLOCK]);
    }
}

// ------------------------------------------------------ LDS-staged walk ---

// Candidates staged per wave per round.  The stage's three planes (x, y, z)
// of kStage floats then lie within 255 dwords of each other, the reach of a
// ds_read2_b32 offset, so one address register serves all three.
constexpr int kStage = 124;

__device__ __forceinline__ int wave_min_i32(int v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d));
    return v;
}

__device__ __forceinline__ int wave_max_i32(int v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d));
    return v;
}

// Orders this wave's LDS accesses on either side: a wave's LDS instructions
// complete in issue order, and the fences keep the compiler from moving
// accesses across this point.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Raw buffer resources (gfx950 word 3: 32-bit format, no swizzle) give
// 32-bit byte offsets with a scalar base, so an element index costs one
// shift instead of 64-bit address arithmetic.  Reads at or past `count`
// elements (the voffset range check) return zeros.  Arrays addressed this way
// hold fewer than 2^28 float4s.
constexpr int kBufWord3 = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t f4_rsrc(const float4* p, uint32_t count)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(p), (short)0, (int)(count * 16u), kBufWord3);
}

// Element voff_elems + soff_elems of the buffer; soff_elems is wave-uniform.
__device__ __forceinline__ float4 f4_load(__amdgpu_buffer_rsrc_t r, uint32_t voff_elems, uint32_t soff_elems = 0u)
{
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, voff_elems * 16u, soff_elems * 16u, 0));
}

// Per-wave LDS words a gather needs: the row table of for_each_row (stride
// 64) or the candidate stage, never both at once.
template <int SUB>
constexpr int gather_lds_words()
{
    return 2 * (2 * SUB + 1) * 64 > 3 * kStage ? 2 * (2 * SUB + 1) * 64 : 3 * kStage;
}

// Candidates in the LDS stage, stored as three planes x[kStage], y[kStage],
// z[kStage].  Lanes reading different candidates then meet in one bank only
// when the candidates are 32 apart (float4 records would collide every 8),
// and two consecutive candidates are one ds_read2 per plane.  Indexing
// yields (x, y, z, 0) and advancing moves along the candidates, as a
// float4 pointer into the slot arrays does.
struct StageRow {
    const float* p;
    __device__ __forceinline__ float4 operator[](uint32_t k) const
    {
        return make_float4(p[k], p[k + kStage], p[k + 2 * kStage], 0.f);
    }
    __device__ __forceinline__ float4 operator*() const { return (*this)[0]; }
    __device__ __forceinline__ StageRow operator+(uint32_t k) const { return {p + k}; }
    __device__ __forceinline__ StageRow& operator+=(uint32_t k)
    {
        p += k;
        return *this;
    }
    __device__ __forceinline__ StageRow& operator++()
    {
        ++p;
        return *this;
    }
};

// Stencil walk of one wave.  When every valid lane of the wave lies in the
// same cell row (cy, cz) - the common case, since slots are cell-sorted x
// fastest - the slot range of each stencil row over the wave's x cells
// widened by SUBX is contiguous.  That range is copied into the wave's LDS
// stage kStage candidates per round, lane k loading slots k, k+64, ...
// (coalesced), and each lane then reads its own pruned sub-range of the round
// from LDS.  With PREFETCH the first round of row k+1 is loaded into
// registers while row k is scanned (8 more VGPRs).  Otherwise every lane walks its own rows through global memory
// (for_each_row).  Both ways visit the same candidates in the same order
// (rows in z, y order, slots ascending), so results do not depend on the way
// taken.  Lanes with valid = false visit nothing but must take part: the call
// is made by the whole wave.
// rP views P (np elements).  body(q, m, j, self_row) visits the m candidates
// q[0..m) whose first slot is j; q is a StageRow in LDS or a float4 pointer
// into P, and q[k] has the candidate's position in x, y, z.  self_row is
// wave-uniform and false only where no lane's own slot can be among the
// candidates: staged rows other than the wave's own cell row.
template <int SUB, bool PRUNE, bool PREFETCH, class Body>
__device__ __forceinline__ void gather_rows(const SceneDev& sc, const uint32_t* __restrict__ start,
                                            const float4* __restrict__ P, __amdgpu_buffer_rsrc_t rP, uint32_t* lds,
                                            bool valid, float fx, float fy, float fz, int cx, int cy, int cz,
                                            Body&& body)
{
    constexpr int kRows = 2 * SUB + 1;
    constexpr int SUBX = kXSplit * SUB;
    constexpr int kPer = (kStage + 63) / 64;
    const int lane = (int)__lane_id();
    const int dx = sc.dims[0], dy = sc.dims[1], dz = sc.dims[2];
    const int rk = cz * dy + cy;
    const int rmin = wave_min_i32(valid ? rk : 0x7fffffff);
    const int rmax = wave_max_i32(valid ? rk : -1);
    if (rmin != rmax) {
        if (valid)
            for_each_row<SUB, PRUNE, 64>(sc, start, fx, fy, fz, cx, cy, cz, lds + lane,
                                         [&](uint32_t b, uint32_t e) { body(P + b, e > b ? e - b : 0u, b, true); });
        return;
    }
    if (rmin == 0x7fffffff) return;  // no valid lane
    // Wave-uniform row and x extent.
    const int row_yz = __builtin_amdgcn_readfirstlane(rmin);
    const int ucy = row_yz % dy, ucz = row_yz / dy;
    const int xlo = __builtin_amdgcn_readfirstlane(max(wave_min_i32(valid ? cx : 0x7fffffff) - SUBX, 0));
    const int xhi = __builtin_amdgcn_readfirstlane(min(wave_max_i32(valid ? cx : -1) + SUBX, dx - 1));
    float* stage = reinterpret_cast<float*>(lds);
    auto fetch = [&](float4 (&dst)[kPer], uint32_t c0, uint32_t n) {
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const uint32_t i = (uint32_t)lane + 64u * (uint32_t)r;
            if (i < n) dst[r] = f4_load(rP, i, c0);
        }
    };
    auto put = [&](const float4 (&src)[kPer], uint32_t n) {
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const uint32_t i = (uint32_t)lane + 64u * (uint32_t)r;
            if (i < n) {
                stage[i] = src[r].x;
                stage[i + kStage] = src[r].y;
                stage[i + 2 * kStage] = src[r].z;
            }
        }
    };
    const int xa = cx > SUBX ? cx - SUBX : 0;
    const int xb = cx < dx - SUBX ? cx + SUBX : dx - 1;
    const int z0 = ucz > SUB ? ucz - SUB : 0, z1 = ucz < dz - SUB ? ucz + SUB : dz - 1;
    constexpr float kReach2 = (float)(SUB * SUB);
    for (int z = z0; z <= z1; ++z) {
        const float gz = z > cz ? (float)z - fz : (z < cz ? fz - (float)(z + 1) : 0.f);
        uint32_t rb[kRows], re[kRows], RB[kRows], RE[kRows];
#pragma unroll
        for (int k = 0; k < kRows; ++k) {
            const int y = ucy - SUB + k;
            rb[k] = re[k] = RB[k] = RE[k] = 0;
            if (y >= 0 && y < dy) {
                const uint32_t row = row_key(sc, y, z);
                RB[k] = start[row + (uint32_t)xlo];
                RE[k] = start[row + (uint32_t)xhi + 1u];
                int x0 = xa, x1 = xb;
                bool keep = valid;
                if constexpr (PRUNE) {
                    const float gy = y > cy ? (float)y - fy : (y < cy ? fy - (float)(y + 1) : 0.f);
                    const float rem = kReach2 - fmaf(gy, gy, gz * gz);
                    keep = keep && rem > 0.f;
                    const float w = (float)kXSplit * __builtin_amdgcn_sqrtf(fmaxf(rem, 0.f));
                    x0 = (int)fmaxf(fminf(floorf(fx - w), (float)(dx - 1)), (float)xa);
                    x1 = (int)fminf(fmaxf(floorf(fx + w), 0.f), (float)xb);
                }
                if (keep) {
                    rb[k] = start[row + (uint32_t)x0];
                    re[k] = start[row + (uint32_t)(x1 + 1)];
                }
            }
        }
        // Slots of row k are [RB[k], RE[k]) with RE[k] >= RB[k] (cell starts
        // are a prefix sum); each lane's [rb[k], re[k]) lies inside.
        float4 pre[kPer];
        fetch(pre, RB[0], min(RE[0] - RB[0], (uint32_t)kStage));
#pragma unroll
        for (int k = 0; k < kRows; ++k) {
            const uint32_t n0 = min(RE[k] - RB[k], (uint32_t)kStage);
            if (!PREFETCH && k > 0) fetch(pre, RB[k], n0);
            put(pre, n0);
            wave_lds_sync();
            if (PREFETCH && k + 1 < kRows) fetch(pre, RB[k + 1], min(RE[k + 1] - RB[k + 1], (uint32_t)kStage));
            // Every valid lane's own slot lies in row (ucy, ucz): k == SUB of z == ucz.
            const bool self_row = z == ucz && k == SUB;
            {
                const uint32_t lo = max(rb[k], RB[k]), hi = min(re[k], RB[k] + n0);
                if (hi > lo) body(StageRow{stage + (lo - RB[k])}, hi - lo, lo, self_row);
            }
            wave_lds_sync();
            // Further rounds of a row longer than the stage.
            for (uint32_t c0 = RB[k] + (uint32_t)kStage; c0 < RE[k]; c0 += (uint32_t)kStage) {
                const uint32_t n = min(RE[k] - c0, (uint32_t)kStage);
                float4 more[kPer];
                fetch(more, c0, n);
                put(more, n);
                wave_lds_sync();
                const uint32_t lo = max(rb[k], c0), hi = min(re[k], c0 + n);
                if (hi > lo) body(StageRow{stage + (lo - c0)}, hi - lo, lo, self_row);
                wave_lds_sync();
            }
        }
    }
}

}  // namespace nprsph
