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
constexpr int kScanItems = 8;
constexpr int kScanBlock = kScanThreads * kScanItems;  // cells per scan block

inline int scan_blocks(uint32_t ncells) { return (int)((ncells + kScanBlock - 1) / kScanBlock); }
// Grids of at most this many scan blocks (8 M cells) skip k_scan_top:
// k_scan_down sums the block totals before each block itself.
constexpr int kScanFoldBlocks = 4096;
inline bool scan_folds(uint32_t ncells) { return scan_blocks(ncells) <= kScanFoldBlocks; }

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
    float* rho;       // rho of slot s (A order); p and 1/rho are PA.w and VA.w
    uint32_t* perm;   // AoS index of slot s
    uint32_t* key;    // cell key of slot s (B order)
    uint32_t* slot;   // arrival rank of slot s within its cell
    uint32_t* torig;  // scatter output: AoS index
    uint32_t* tsrc;   // scatter output: B slot
    uint32_t* tkey;   // scatter output: cell key
    uint32_t* count;  // per-cell counts (all zero between builds)
    uint32_t* start;  // per-cell first slot, total_cells+1 entries
    uint32_t* bsum;   // scan block sums (counted by the fused step, else summed by k_scan_reduce)
    double* partials; // reduction scratch, 8 doubles per tile
};

bool tuning_supported(const sph_tuning_t& t);

// ---- learned coarse step -------------------------------------------------

constexpr int kMlpIn = 9, kMlpHidden = 64, kMlpOut = 3;
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpHidden * kMlpOut + kMlpOut;  // 4995

// GELU as 0.5x + q(min(x^2, 9)) + 0.5*max(|x|-3, 0), q(s) = s(c0 + s(c1 + s(c2 + s c3))):
// a transcendental-free form of x*Phi(x), max |error| 2.8e-3.
constexpr float kGeluC0 = 3.88549402e-01f, kGeluC1 = -5.37665106e-02f, kGeluC2 = 4.76835308e-03f,
                kGeluC3 = -1.70698904e-04f, kGeluClamp = 9.0f, kGeluKnee = 3.0f;

struct MlpNorm {
    float v_scale;    // velocity difference scale
    float rho_scale;  // density scale
    float pi_scale;   // pressure-term scale
    float acc_scale;  // acceleration per unit of summed message
};

// Device buffers of the MLP (parameters, optimiser state) and of the current
// materialised pair batch.  Grown on demand; pair buffers hold `pair_cap`
// rows, target buffers `target_cap` entries.
struct MlpWork {
    MlpNorm norm;
    float* params = nullptr;  // kMlpParams float32 master weights
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* grad = nullptr;
    uint64_t pair_cap = 0, target_cap = 0, partial_cap = 0;
    uint32_t* cnt = nullptr;   // per target
    uint64_t* off = nullptr;   // per target + 1
    float* tgt = nullptr;      // 3 per target
    float* wt = nullptr;       // per target
    float* pred = nullptr;     // 3 per target (or per slot for the coarse step)
    float* dpred = nullptr;    // 3 per target
    double* lossp = nullptr;   // per target
    uint32_t* pj = nullptr;    // per pair: neighbour slot
    uint32_t* pt = nullptr;    // per pair: target index
    float* X = nullptr;        // 9 per pair
    float* Z1 = nullptr;       // 64 per pair
    float* Z2 = nullptr;
    float* dZ1 = nullptr;
    float* dZ2 = nullptr;
    float* M = nullptr;        // 3 per pair
    float* partial = nullptr;  // kMlpParams per gradient block
    __bf16* pack_a = nullptr;  // MFMA operand images (k_mlp_pack)
    float* pack_c = nullptr;
    uint16_t* gelu_tab = nullptr;  // BF16 GELU table of the MFMA kernel (k_gelu_table)
    unsigned long long* pairs = nullptr;  // pairs evaluated by the last MFMA message pass
    bool pack_stale = true;    // params changed since the last pack
};

constexpr int kMlpPackA = 4 * 64 * 8 + 4 * 2 * 64 * 8 + 2 * 64 * 8;  // bf16 entries
constexpr int kMlpPackC = 4 * 64 * 4 * 2 + 64 * 4;                  // float entries
constexpr int kMlpGeluTab = 4096;                                    // GELU table entries

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
// Pass: force_comp.glsl alone (writes force into the AoS).  FusedAos / Fused:
// force and integrate in one kernel, with or without the AoS record write.
enum class ForceMode : int { Pass = 0, FusedAos = 1, Fused = 2 };
hipError_t launch_force(hipStream_t s, const DeviceState& d, ForceMode mode);
hipError_t launch_integrate_aos(hipStream_t s, const DeviceState& d);
// Per-tile partials of tiles [first_tile, first_tile+ntiles) into partials[8*ntiles];
// the fold combines them into out[7] in tile order.
hipError_t launch_metrics(hipStream_t s, const DeviceState& d, int first_tile, int ntiles, double* partials);
hipError_t launch_metrics_fold(hipStream_t s, const double* partials, int ntiles, double* out);

hipError_t launch_pairs(hipStream_t s, const DeviceState& d, int nscenes, const MlpWork& w, uint32_t first,
                        uint32_t stride, uint32_t T, bool emit);
hipError_t launch_mlp_forward(hipStream_t s, const MlpWork& w, uint64_t P);
hipError_t launch_mlp_aggregate(hipStream_t s, const MlpWork& w, uint32_t T, float inv_wsum);
hipError_t launch_mlp_backward(hipStream_t s, const MlpWork& w, uint64_t P, uint32_t rows_per_block, uint32_t nblocks);
hipError_t launch_mlp_adam(hipStream_t s, const MlpWork& w, float lr, float beta1, float beta2, float eps, float bc1,
                           float bc2);
hipError_t launch_train_targets(hipStream_t s, const DeviceState& d, int nscenes, const MlpWork& w,
                                const sph_particle_t* later, uint32_t first, uint32_t stride, uint32_t T);
hipError_t launch_coarse_integrate(hipStream_t s, const DeviceState& d, const MlpWork& w, bool use_force);
hipError_t launch_mlp_pack(hipStream_t s, const MlpWork& w);
hipError_t launch_coarse_mlp(hipStream_t s, const DeviceState& d, int nscenes, const MlpWork& w, int path);
hipError_t launch_pred_to_aos(hipStream_t s, const DeviceState& d, const MlpWork& w, float* out);

}  // namespace nprsph
