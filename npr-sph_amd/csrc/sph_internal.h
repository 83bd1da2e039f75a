// sph_internal.h - host-side launchers of the solver kernels (sph_kernels.hip)
// used by the C ABI implementation (sph_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph.h"
#include "sph_device.h"

namespace nprsph {

// Work layout of the exclusive scan over cell counts.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanBlock = kScanThreads * kScanItems;  // cells per scan block

inline int scan_blocks(uint32_t ncells) { return (int)((ncells + kScanBlock - 1) / kScanBlock); }

// Device buffers of one solver instance.  "A" is the cell-sorted state the
// gathers read; "B" is the integrated state in the same slot order, from which
// the next neighbour build starts.
struct DeviceState {
    const SceneDev* scenes;
    const Tile* tiles;   // kTile-particle tiles (grid build, streaming passes)
    int ntiles;
    const Tile* gtiles;  // tiles of tuning.block particles (gathers)
    int ngtiles;
    sph_tuning_t tuning;
    uint32_t n;
    uint32_t total_cells;
    sph_particle_t* aos;
    float4* PA;
    float4* VA;
    float4* PB;
    float4* VB;
    float* rho;       // rho of slot s (A order)
    float2* D2;       // (p, 1/rho) of slot s
    uint32_t* perm;   // AoS index of slot s
    uint32_t* key;    // cell key of slot s (B order)
    uint32_t* slot;   // arrival rank of slot s within its cell
    uint32_t* torig;  // scatter output: AoS index
    uint32_t* tsrc;   // scatter output: B slot
    uint32_t* tkey;   // scatter output: cell key
    uint32_t* count;  // per-cell counts (all zero between builds)
    uint32_t* start;  // per-cell first slot, total_cells+1 entries
    uint32_t* bsum;   // scan block sums
    double* partials; // reduction scratch, 8 doubles per tile
};

bool tuning_supported(const sph_tuning_t& t);

// Each launcher enqueues one kernel on `s` and returns hipGetLastError().
hipError_t launch_set_scene(hipStream_t s, SceneDev* dst, const SceneDev& value);
hipError_t launch_fill_aos_defaults(hipStream_t s, sph_particle_t* aos, uint32_t n);
hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t n, uint32_t value);
hipError_t launch_identity(hipStream_t s, uint32_t* p, uint32_t n);
hipError_t launch_lattice(hipStream_t s, sph_particle_t* aos, uint32_t first, const sph_lattice_t& spec);
hipError_t launch_refresh(hipStream_t s, const DeviceState& d, bool pos_vel, bool extras);
hipError_t launch_hash_count(hipStream_t s, const DeviceState& d);
hipError_t launch_scan_reduce(hipStream_t s, const DeviceState& d);
hipError_t launch_scan_top(hipStream_t s, const DeviceState& d);
hipError_t launch_scan_down(hipStream_t s, const DeviceState& d);
hipError_t launch_scatter(hipStream_t s, const DeviceState& d);
hipError_t launch_rank_reorder(hipStream_t s, const DeviceState& d);
hipError_t launch_density(hipStream_t s, const DeviceState& d, bool write_aos);
hipError_t launch_force(hipStream_t s, const DeviceState& d, bool fuse_integrate);
hipError_t launch_integrate_aos(hipStream_t s, const DeviceState& d);
hipError_t launch_metrics(hipStream_t s, const DeviceState& d, int first_tile, int ntiles);

}  // namespace nprsph
