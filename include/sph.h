/*
 * sph.h - C ABI of the MI355X SPH solver (libnprsph.so).
 *
 * Drop-in replacement for the three GL compute programs of NPR-SPH
 * (rho_pres_comp.glsl, force_comp.glsl, integrate_comp.glsl) and the host
 * code that owns their buffers and dispatches them (NPR-SPH/Main.cpp).
 * Plain C: no HIP, torch or C++ types cross this boundary.  Streams are
 * passed as `void*` (a hipStream_t, or NULL to let the library create one).
 *
 * Conventions (mirroring NPR-SPH/InitShader.cpp:47-112 and
 * NPR-SPH/Main.cpp:428-451, where a failed shader build returns (GLuint)-1,
 * logs the reason and leaves the previous program in use):
 *   - every int-returning call returns SPH_OK (0) or a negative SPH_ERR_*;
 *   - on failure the context stays valid and sph_last_error() holds the
 *     reason (sph_create_error() for failures of the create calls);
 *   - calls are asynchronous on the context's stream; only sph_download,
 *     sph_download_grid, sph_metrics and sph_synchronize wait for it;
 *   - one context per stream; a context is not thread-safe.
 */
#ifndef NPR_SPH_AMD_SPH_H
#define NPR_SPH_AMD_SPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_ABI_VERSION 1

enum {
    SPH_OK = 0,
    SPH_ERR_INVALID = -1, /* bad argument: null pointer, size, scene index, non-finite constant */
    SPH_ERR_HIP = -2,     /* a HIP runtime call failed; message holds hipGetErrorString */
    SPH_ERR_NOMEM = -3,   /* device or host allocation failed */
    SPH_ERR_STATE = -4    /* call not valid in the context's current state */
};

/* One particle, 64 bytes, std430 layout of the SSBO at binding 0.
 * Replaces `struct Particle` of NPR-SPH/Main.cpp:93-99 and
 * rho_pres_comp.glsl:12-18 / force_comp.glsl:12-18 / integrate_comp.glsl:14-20.
 * extras = (rho, pressure, age, unused); pos.w, vel.w, force.w, extras.z and
 * extras.w are never written by the solver. */
typedef struct sph_particle {
    float pos[4];
    float vel[4];
    float force[4];
    float extras[4];
} sph_particle_t;

/* ConstantsUniform, NPR-SPH/Main.cpp:110-116 (std140 block at binding 1,
 * rho_pres_comp.glsl:25-31).  Defaults: {0.02, 4.0, 3000, 1000}. */
typedef struct sph_constants {
    float mass;
    float smoothing_coeff;
    float visc;
    float resting_rho;
} sph_constants_t;

/* BoundaryUniform, NPR-SPH/Main.cpp:118-122 (std140 block at binding 2,
 * integrate_comp.glsl:27-31).  Defaults: upper (0.5,1,0.5,1), lower
 * (-0.1,-0.35,-0.1,1).  The .w lanes are ignored. */
typedef struct sph_boundary {
    float upper[4];
    float lower[4];
} sph_boundary_t;

/* Constants the reference bakes into its shaders:
 * PARTICLE_RADIUS 0.005 (rho_pres_comp.glsl:5), GAS_CONST 2000
 * (rho_pres_comp.glsl:33), G = (0,-9806.65,0) (force_comp.glsl:33),
 * DAMPING 0.3 (integrate_comp.glsl:8), dt = 1/NUM_PARTICLES = 1e-4
 * (integrate_comp.glsl:33; kept at 1e-4 for every particle count). */
typedef struct sph_fixed {
    float particle_radius;
    float gas_const;
    float gravity[3];
    float damping;
    float dt;
} sph_fixed_t;

/* Seeded jittered lattice (device-side scene initialiser, replaces
 * make_grid/init_particles of NPR-SPH/Main.cpp:488-539).  Particle
 * k = (ix*ny + iy)*nz + iz (x outermost, z innermost, as make_grid) sits at
 * origin + i*spacing + jitter, jitter per axis uniform in [-jitter, +jitter)
 * from a counter-based SplitMix64 stream keyed by (seed, 3k+axis).
 * Only the first `count` lattice sites are used (count <= nx*ny*nz).
 * vel = force = extras = 0, pos.w = 1. */
typedef struct sph_lattice {
    int32_t nx, ny, nz;
    uint32_t count;
    float origin[3];
    float spacing;
    float jitter;
    uint64_t seed;
} sph_lattice_t;

/* Cells are SPH_GRID_X_SPLIT times narrower along x than along y and z. */
#define SPH_GRID_X_SPLIT 2
/* Cells are keyed in bands of SPH_GRID_BAND z-planes (see sph_grid_info_t.ncells). */
#define SPH_GRID_BAND 8

/* Geometry of one scene's uniform neighbour grid (read-only diagnostics).
 * Cell coordinate along an axis = clamp(floor((p - lo) * inv), 0, dims - 1) in
 * float32, inv = SPH_GRID_X_SPLIT * inv_cell (exact) along x, inv_cell along
 * y and z. */
typedef struct sph_grid_info {
    float lo[3];          /* grid origin = boundary lower corner */
    float inv_cell;       /* 1 / cell edge along y and z (float32, used by the cell hash) */
    double cell;          /* cell edge along y and z (>= smoothing length / cells_per_h);
                             cell / SPH_GRID_X_SPLIT along x */
    int32_t dims[3];      /* cells per axis */
    uint32_t ncells;      /* dims[0]*dims[1]*B*ceil(dims[2]/B), B = SPH_GRID_BAND: cells are keyed
                             band by band, key = ((z/B*dims[1] + y)*B + z%B)*dims[0] + x;
                             cell `ncells` holds NaN positions */
    uint32_t cell_base;   /* global index of this scene's first cell */
    uint32_t first;       /* first particle of this scene in the AoS buffer */
    uint32_t count;       /* particles in this scene */
    int32_t cells_per_h;  /* stencil reach in cells */
} sph_grid_info_t;

/* Launch and grid parameters of the density/force gathers.  Every setting
 * gives results within the FP32 tolerance of the FP64 transcription; a fixed
 * setting gives bit-identical results run to run, and prune/block/
 * blocks_per_xcd never change the bits (only cells_per_h changes the
 * neighbour summation order).  Defaults: {2, 1, 128, 0}. */
typedef struct sph_tuning {
    int32_t cells_per_h;     /* 1..3: cell edge = h*(1+2^-10)/cells_per_h along y and z,
                                SPH_GRID_X_SPLIT times narrower along x */
    int32_t prune;           /* 1: skip stencil rows and cells farther than h from the particle */
    int32_t block;           /* 128 or 256 particles per gather workgroup */
    int32_t blocks_per_xcd;  /* 0: one workgroup per tile; >0: persistent workgroups per XCD */
} sph_tuning_t;

/* Per-scene reductions over the current state (sph_metrics).  Kinetic
 * energy, centre of mass and max speed are taken over the particles whose
 * pos and vel are finite. */
typedef struct sph_metrics {
    double mean_rho;      /* mean of extras.x over all particles */
    double kinetic;       /* sum of 0.5*mass*|vel|^2 */
    double com[3];        /* mean position */
    double max_speed;     /* max |vel| */
    uint64_t nonfinite;   /* particles with a non-finite pos or vel component */
} sph_metrics_t;

typedef struct sph_ctx sph_ctx;

/* Reference defaults (Main.cpp:110-122 and the shader-baked constants). */
void sph_defaults(sph_constants_t* c, sph_boundary_t* b, sph_fixed_t* f);

/* Create a single-scene solver for n particles (1 <= n <= 2^28 - 2, the total
 * for a batch as well; the slot arrays are addressed with 32-bit byte
 * offsets).  Any of c/b/f may be
 * NULL for the defaults.  The AoS buffer starts zeroed except pos.w = 1.
 * Replaces initOpenGL's buffer setup (Main.cpp:524-527, 632-641). */
int sph_create(sph_ctx** out, size_t n, const sph_constants_t* c, const sph_boundary_t* b,
               const sph_fixed_t* f, void* stream);

/* Create a solver holding `nscenes` independent scenes of counts[s] particles,
 * stored back to back in one AoS buffer (scene s starts at sum(counts[<s])).
 * Every pass advances all scenes in one launch.  c/b/f are arrays of
 * nscenes entries or NULL for the defaults.  No reference counterpart: the
 * reference runs one scene. */
int sph_create_batch(sph_ctx** out, int nscenes, const size_t* counts, const sph_constants_t* c,
                     const sph_boundary_t* b, const sph_fixed_t* f, void* stream);

/* Message of the last failed sph_create / sph_create_batch on this thread. */
const char* sph_create_error(void);

void sph_destroy(sph_ctx* ctx);
const char* sph_last_error(const sph_ctx* ctx);

size_t sph_size(const sph_ctx* ctx);     /* total particles */
int sph_num_scenes(const sph_ctx* ctx);

/* Parameter upload (sendUniforms, Main.cpp:254-282).  Takes effect for every
 * pass enqueued after the call; scene = -1 applies to all scenes.  A change of
 * smoothing length or box that needs a different grid is applied here too. */
int sph_set_constants(sph_ctx* ctx, int scene, const sph_constants_t* c);
int sph_set_boundary(sph_ctx* ctx, int scene, const sph_boundary_t* b);
int sph_set_fixed(sph_ctx* ctx, int scene, const sph_fixed_t* f);
int sph_get_params(const sph_ctx* ctx, int scene, sph_constants_t* c, sph_boundary_t* b, sph_fixed_t* f);

/* Host <-> device copies of particles [first, first+n) of the AoS buffer
 * (glBufferData of init_particles, Main.cpp:526).  sph_download waits. */
int sph_upload(sph_ctx* ctx, const sph_particle_t* host, size_t first, size_t n);
int sph_download(sph_ctx* ctx, sph_particle_t* host, size_t first, size_t n);

/* Device-side scene initialiser: fills scene `scene` from a lattice spec. */
int sph_init_lattice(sph_ctx* ctx, int scene, const sph_lattice_t* spec);

/* The three-pass surface (Main.cpp:295-303).  Each pass reads and writes the
 * AoS buffer exactly as its shader does:
 *   sph_rho_pres : extras.x = rho, extras.y = p     (rho_pres_comp.glsl:35-58)
 *   sph_force    : force.xyz                        (force_comp.glsl:35-66)
 *   sph_integrate: vel.xyz, pos.xyz                 (integrate_comp.glsl:35-81)
 * Stream order replaces glMemoryBarrier. */
int sph_rho_pres(sph_ctx* ctx);
int sph_force(sph_ctx* ctx);
int sph_integrate(sph_ctx* ctx);

/* nsteps x (rho_pres, force, integrate), the per-frame block of
 * Main.cpp:291-304, run as captured HIP graphs on cell-sorted state.  The AoS
 * buffer afterwards equals what the three passes would leave, bit for bit.
 * Steps that a later step of the same call follows may skip the AoS write
 * and work on the sorted state alone; the buffer is complete once the call's
 * work on the stream has finished.
 * nsteps = 0 advances nothing: it prepares the sorted state and captures the
 * step graphs (16-step and 1-step), so that a later call launches only.
 * At most 8 graph launches (16 steps each, or 1 for the remainder) are queued
 * on the stream at a time: a call for more steps waits on the host for the
 * oldest queued launch before it adds the next. */
int sph_step(sph_ctx* ctx, int nsteps);

/* Device pointer of the AoS buffer (original index order, stride 64 bytes, pos
 * at offset 0: the vertex attribute a renderer binds, Main.cpp:533-535). */
int sph_device_buffer(sph_ctx* ctx, sph_particle_t** d_aos, size_t* n);

/* The context's stream (a hipStream_t). */
void* sph_stream(sph_ctx* ctx);
int sph_synchronize(sph_ctx* ctx);

/* Diagnostics: grid geometry of a scene; the neighbour grid built by the last
 * rho_pres/step (cell_start has total_cells+1 entries over all scenes,
 * perm[slot] = AoS index of the particle at cell-sorted slot). */
int sph_grid_info(const sph_ctx* ctx, int scene, sph_grid_info_t* out);
size_t sph_total_cells(const sph_ctx* ctx);
int sph_download_grid(sph_ctx* ctx, uint32_t* cell_start, size_t ncell_entries, uint32_t* perm, size_t n);

/* On-disk snapshot of the particle buffer (no reference counterpart; the
 * reference keeps its state in the GL buffer only).  The file is a 64-byte
 * little-endian header:
 *   0  char[8]  "NPRSPHv1"       8  uint32 header bytes (64)
 *   12 uint32   record bytes (64) 16 uint64 particle count
 *   24 uint32   scene count       28 uint32 reserved (0)
 *   32 uint64   frame (the caller's tag)   40..63 reserved (0)
 * then one uint64 particle count per scene, then the sph_particle_t records
 * in original index order: pos at offset 0, stride 64, the layout a renderer
 * binds (Main.cpp:533-535).  sph_load reads a file with the same per-scene
 * counts and replaces the state (as sph_upload); frame may be NULL. */
int sph_save(sph_ctx* ctx, const char* path, uint64_t frame);
int sph_load(sph_ctx* ctx, const char* path, uint64_t* frame);

/* Reductions over scene `scene` of the AoS state (waits). */
int sph_metrics(sph_ctx* ctx, int scene, sph_metrics_t* out);

/* sph_metrics without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate (serving loops, harnesses that forbid
 * hipMalloc): it calls no hipMalloc, hipFree or allocating library routine
 * and keeps its partial sums in a static device array of the library (1 MiB
 * per GPU, held until the process ends).  Results equal sph_metrics bit for
 * bit.  Limit: scenes of at most 4,194,304 particles (16,384 tiles of 256);
 * larger scenes return SPH_ERR_INVALID.  All calls on one GPU share the array:
 * calls from contexts on different streams of the same GPU that overlap in
 * time race on it.  Waits for the context's stream. */
int sph_metrics_noalloc(sph_ctx* ctx, int scene, sph_metrics_t* out);

/* Gather tuning (no reference counterpart).  sph_autotune times every
 * supported setting for `trial_steps` steps from the current state, re-times
 * the four fastest and the setting in force for 4*trial_steps steps, restores
 * that state, and applies the fastest of those if it beats the setting in
 * force by more than 2% (otherwise keeps the setting in force).  It reports
 * the applied setting with its per-step milliseconds. */
int sph_get_tuning(const sph_ctx* ctx, sph_tuning_t* out);
int sph_set_tuning(sph_ctx* ctx, const sph_tuning_t* t);
int sph_autotune(sph_ctx* ctx, int trial_steps, sph_tuning_t* best, double* best_ms_per_step);

/* ---- Learned coarse step (no reference counterpart) ----------------------
 * A pair-message MLP over every neighbour pair within h: features
 *   (x_i-x_j)/h, (v_i-v_j)*v_scale, rho_i*rho_scale, rho_j*rho_scale,
 *   (p_i+p_j)/(2 rho_j)*pi_scale
 * -> 64 -> GELU -> 64 -> GELU -> 3, summed per particle: a_i = acc_scale*sum + G.
 * One coarse step replaces 4 fine steps: v' = v + 4dt a, x' = x + 4dt v +
 * 10 dt^2 a (semi-implicit Euler with a held fixed), then the wall clamp.
 * Parameters: float32, SPH_MLP_PARAMS values laid out as
 * W1[9][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3]. */
#define SPH_MLP_PARAMS 4995

typedef struct sph_mlp_norm {
    float v_scale;
    float rho_scale;
    float pi_scale;
    float acc_scale;
} sph_mlp_norm_t;

typedef struct sph_train_opts {
    uint32_t first;    /* target particles: cell-sorted slots first, first+stride, ... */
    uint32_t stride;
    float lr;          /* Adam step size and moments */
    float beta1;
    float beta2;
    float eps;
    int32_t update;    /* 0: gradient only; 1: gradient and one Adam step */
} sph_train_opts_t;

typedef struct sph_train_result {
    double loss;       /* weighted mean squared error of the summed messages */
    uint64_t pairs;    /* neighbour pairs in the batch */
    uint32_t targets;
    uint32_t weighted; /* targets with weight 1 (off the walls, finite) */
} sph_train_result_t;

typedef struct sph_coarse_opts {
    int32_t use_network;  /* 1: learned messages; 0: the SPH force held over the 4 fine steps */
    int32_t judge_every;  /* every judge_every-th coarse step is also run as 4 plain steps (0: never) */
    float drift_limit;    /* judge: relative drift |mean rho(coarse) - mean rho(plain)| / mean rho(plain)
                             from the same state; above this the plain result is kept (0.01 = 1%) */
    int32_t cooldown;     /* after a rejected step, this many coarse steps run as 4 plain steps each,
                             and the next step is judged again (another rejection restarts the
                             cooldown); 0: only the judged step is replaced */
} sph_coarse_opts_t;

typedef struct sph_coarse_stats {
    int64_t coarse_steps;  /* accepted coarse steps */
    int64_t plain_steps;   /* fine steps run by the plain solver (fallback and cooldown) */
    int64_t judged;
    int64_t fallbacks;
    double last_drift;     /* relative drift of the last judged step */
    double max_drift;
    int64_t pairs;         /* neighbour pairs the network evaluated in the last message pass */
} sph_coarse_stats_t;

/* Load / read the network (float32 master weights) and its input scales. */
int sph_mlp_set(sph_ctx* ctx, const float* params, size_t n, const sph_mlp_norm_t* norm);
int sph_mlp_get(sph_ctx* ctx, float* params, size_t n, sph_mlp_norm_t* norm);

/* One training step on the GPU.  `ctx` holds the state at time t, `later` the
 * same particles 4 fine steps later (equal particle counts and scenes).
 * Targets are the mean accelerations (v(t+4dt)-v(t))/(4dt) - G of the chosen
 * particles in units of acc_scale; particles on a wall or non-finite get
 * weight 0.  Runs the density pass on ctx, enumerates the neighbour pairs,
 * runs the forward and backward kernels and, if opts->update, one Adam step.
 * grad_out (may be NULL) receives the SPH_MLP_PARAMS gradient. */
int sph_mlp_train(sph_ctx* ctx, const sph_ctx* later, const sph_train_opts_t* opts, float* grad_out,
                  sph_train_result_t* result);

/* Summed messages of every particle at the current state (AoS order, 3 floats
 * per particle, units of acc_scale).  path 0: float32 kernel; 1: BF16 MFMA
 * kernel.  Runs the density pass first (as sph_rho_pres). */
int sph_mlp_predict(sph_ctx* ctx, int path, float* out);

/* n coarse steps (each = 4 fine steps of simulated time), with the plain
 * solver as judge and fallback.  stats may be NULL. */
int sph_coarse_step(sph_ctx* ctx, int n, const sph_coarse_opts_t* opts, sph_coarse_stats_t* stats);

/* Copy the whole AoS state of src into dst (same particle count). */
int sph_copy_state(sph_ctx* dst, const sph_ctx* src);

/* Kernel timing: while enabled, passes run without graphs and every kernel is
 * bracketed by HIP events on the context's stream.  sph_kernel_times returns
 * the accumulated milliseconds and launch counts per kernel name. */
int sph_set_timing(sph_ctx* ctx, int enabled);
int sph_kernel_times(sph_ctx* ctx, int max_entries, char (*names)[32], double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* NPR_SPH_AMD_SPH_H */
