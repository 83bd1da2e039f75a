// sph_device.h - device-side data layout and per-particle math shared by the
// solver kernels.  Everything here is compiled for gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph.h"

namespace nprsph {

// Particles per tile of the particle-parallel kernels (one workgroup, four
// wave64s).  Tiles never straddle scenes.
constexpr int kTile = 256;

// Per-scene parameter block: the device copy of ConstantsUniform (binding 1),
// BoundaryUniform (binding 2) and the shader-baked constants, plus the scene's
// grid geometry and particle range.  Kernels read it with scalar loads, so a
// parameter change is one small device write and needs no graph re-capture.
struct alignas(16) SceneDev {
    float lo[3];       // grid origin (boundary lower corner)
    float inv_cs;      // 1 / cell edge along y and z, float32 as the hash uses it
    int32_t dims[3];   // cells per axis
    uint32_t cell_base;  // global index of the scene's first cell
    uint32_t ncells;   // dx*dy*kCellBand*ceil(dz/kCellBand); cell cell_base+ncells holds NaN positions
    uint32_t p_begin;  // first particle (AoS index == cell-sorted slot range)
    uint32_t p_end;
    float inv_csx;     // 1 / cell edge along x = kXSplit * inv_cs (exact)
    float h;           // smoothing length, smoothing_coeff*PARTICLE_RADIUS (rho_pres_comp.glsl:40)
    float h2;
    float poly6_m;     // mass*315/(64*pi*h^9)                 (rho_pres_comp.glsl:52)
    float rho0;        // resting_rho
    float gas_const;   // GAS_CONST                            (rho_pres_comp.glsl:33)
    float pres_m;      // mass*45/(pi*h^6)/2  (pressure term)  (force_comp.glsl:41,59)
    float visc_m;      // mass*45/(pi*h^6)*visc (viscosity)    (force_comp.glsl:42,60,63)
    float dt;          // time step                            (integrate_comp.glsl:33)
    float grav[3];     // G                                    (force_comp.glsl:33)
    float damping;     // DAMPING                              (integrate_comp.glsl:8)
    float lower[3];    // box lower corner                     (integrate_comp.glsl:27-31)
    float mass;
    float upper[3];    // box upper corner
    float pad1;
};
static_assert(sizeof(SceneDev) == 128, "SceneDev layout");

// Cells are ordered by bands of kCellBand z-planes: key of cell (x, y, z) =
// cell_base + ((band*dy + y)*kCellBand + z%kCellBand)*dx + x, band = z/kCellBand.
// A stencil row (fixed y, z) is still a contiguous run of cells along x.  The
// cell-sorted slots sweep the domain band by band, each band bottom to top,
// so the gathers' working window (the rows within reach inside one band) stays
// small enough for an XCD's L2 on large scenes, where whole z-planes would
// not; and the eighths of the slots that the XCDs take are still slabs of z,
// each holding full columns of the fluid's depth.  The last band of a scene
// may be partial; its missing planes are cells that stay empty.
constexpr int kCellBand = SPH_GRID_BAND;

// Cells are kXSplit times narrower along x than along y and z.  A stencil row
// (fixed y, z) is a contiguous slot range whatever the x edge, so narrower x
// cells trim the chord ends of each row (fewer candidates) without adding
// rows; only the cell count grows.  A particle's stencil reaches kXSplit*sub
// cells along x and sub cells along y and z.
constexpr int kXSplit = SPH_GRID_X_SPLIT;

struct alignas(16) Tile {
    uint32_t scene;
    uint32_t begin;  // first particle / slot of the tile
    uint32_t end;    // one past the last
    uint32_t pad;
};

// Cell coordinates of a position: c = clamp(floor((x-lo)*inv), 0, dim-1), inv
// = inv_csx along x and inv_cs along y and z.  Float32 with no contraction, so
// the CPU build reproduces it bit for bit.  Returns false when any component is
// NaN (such particles go to the scene's NaN cell, which is in no stencil: every
// comparison with NaN is false in the reference as well).  With edges >= h/sub
// (y, z) and h/(kXSplit*sub) (x), clamping keeps |xi-xj| < h => |ci-cj| <= sub
// along y and z and <= kXSplit*sub along x.
// f* receive the unclamped fractional cell coordinates.
__device__ __forceinline__ bool cell_coords(const SceneDev& sc, float x, float y, float z, float& fx, float& fy,
                                            float& fz, int& cx, int& cy, int& cz)
{
#pragma clang fp contract(off)
    fx = (x - sc.lo[0]) * sc.inv_csx;
    fy = (y - sc.lo[1]) * sc.inv_cs;
    fz = (z - sc.lo[2]) * sc.inv_cs;
    if (fx != fx || fy != fy || fz != fz) return false;
    cx = fx < 0.0f ? 0 : (fx >= (float)sc.dims[0] ? sc.dims[0] - 1 : (int)fx);
    cy = fy < 0.0f ? 0 : (fy >= (float)sc.dims[1] ? sc.dims[1] - 1 : (int)fy);
    cz = fz < 0.0f ? 0 : (fz >= (float)sc.dims[2] ? sc.dims[2] - 1 : (int)fz);
    return true;
}

// Key of cell (0, y, z): the first cell of stencil row (y, z).
__device__ __forceinline__ uint32_t row_key(const SceneDev& sc, int y, int z)
{
    const uint32_t band = (uint32_t)z / (uint32_t)kCellBand, zb = (uint32_t)z % (uint32_t)kCellBand;
    return sc.cell_base + ((band * (uint32_t)sc.dims[1] + (uint32_t)y) * (uint32_t)kCellBand + zb) * (uint32_t)sc.dims[0];
}

__device__ __forceinline__ uint32_t global_key(const SceneDev& sc, float x, float y, float z)
{
    int cx, cy, cz;
    float fx, fy, fz;
    if (!cell_coords(sc, x, y, z, fx, fy, fz, cx, cy, cz)) return sc.cell_base + sc.ncells;
    return row_key(sc, cy, cz) + (uint32_t)cx;
}

// integrate_comp.glsl:41-77 for one particle: a = F/rho, v' = v + dt*a,
// x' = x + dt*v', then per axis the lower wall is tested before the upper
// wall and a hit clamps the position and scales the velocity by -DAMPING.
// No contraction: the fused step and the stand-alone pass produce the same bits.
__device__ __forceinline__ void integrate_one(const SceneDev& sc, const float x[3], const float v[3],
                                              const float f[3], float rho, float nx[3], float nv[3])
{
#pragma clang fp contract(off)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float acc = f[a] / rho;
        float vel = v[a] + sc.dt * acc;
        float pos = x[a] + sc.dt * vel;
        if (pos < sc.lower[a]) {
            pos = sc.lower[a];
            vel *= -sc.damping;
        } else if (pos > sc.upper[a]) {
            pos = sc.upper[a];
            vel *= -sc.damping;
        }
        nx[a] = pos;
        nv[a] = vel;
    }
}

// SplitMix64 output function (Steele, Lea and Flood, OOPSLA 2014) used as a
// counter-based stream: value(seed, k) = mix(seed + (k+1)*golden).
__device__ __forceinline__ uint64_t splitmix64(uint64_t seed, uint64_t counter)
{
    uint64_t z = seed + (counter + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Jitter in [-amp, amp): top 24 bits centred, times amp*2^-23.
__device__ __forceinline__ float lattice_jitter(uint64_t seed, uint64_t counter, float amp)
{
#pragma clang fp contract(off)
    const int32_t u = (int32_t)(splitmix64(seed, counter) >> 40) - (1 << 23);
    const float scale = amp * (1.0f / 8388608.0f);
    return (float)u * scale;
}

}  // namespace nprsph
