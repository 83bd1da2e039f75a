// sph_gather.h - neighbour enumeration shared by the gather kernels
// (density, force) and the learned coarse step's pair kernels.
#pragma once

#include "sph_device.h"

namespace nprsph {

// Visits the stencil rows of a particle at fractional cell coordinates
// (fx,fy,fz) in cell (cx,cy,cz): rows (y,z) within SUB cells, each a slot
// range [b,e) covering x-adjacent cells.  With PRUNE, rows whose y/z gap to
// the particle is >= SUB cells are skipped and the x extent is cut to
// floor(fx -+ w), w = sqrt(SUB^2 - gap^2).  SUB cells exceed h by the factor
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
    const int dx = sc.dims[0], dy = sc.dims[1], dz = sc.dims[2];
    const int xa = cx > SUB ? cx - SUB : 0;
    const int xb = cx < dx - SUB ? cx + SUB : dx - 1;
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
                    const float w = __builtin_amdgcn_sqrtf(fmaxf(rem, 0.f));
                    // Cells 0 and dx-1 extend to -inf/+inf, so the window ends
                    // are clamped to them before meeting the stencil bounds.
                    x0 = (int)fmaxf(fminf(floorf(fx - w), (float)(dx - 1)), (float)xa);
                    x1 = (int)fminf(fmaxf(floorf(fx + w), 0.f), (float)xb);
                }
                if (keep) {
                    const uint32_t row = sc.cell_base + ((uint32_t)z * (uint32_t)dy + (uint32_t)y) * (uint32_t)dx;
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

}  // namespace nprsph
