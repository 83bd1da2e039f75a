// sph_mlp.hip - kernels of the learned coarse step (no reference counterpart).
//
// Network: a pair-message MLP 9 -> 64 -> GELU -> 64 -> GELU -> 3 over every
// neighbour pair within h.  Messages are summed per particle and scaled to an
// acceleration; one coarse step advances 4 fine steps.  Parameter layout
// (float32, 4995 values): W1[9][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3].
//
// Training runs on materialised pair batches in FP32:
//   k_pairs_count / k_pairs_emit  neighbour pairs + features of a target set
//   k_mlp_forward_rows            z1, z2 and messages per pair
//   k_mlp_aggregate               per-target sums, loss terms and dL/dpred
//   k_mlp_backward_rows           dz2, dz1 per pair
//   k_mlp_wgrad_partial/_reduce   weight gradients as block partial sums
//                                 added in a fixed order (deterministic)
//   k_mlp_adam                    Adam update of the float32 master weights
// The coarse step integrates with k_coarse_integrate.
#include <algorithm>

#include "sph_gather.h"
#include "sph_internal.h"

namespace nprsph {

namespace {

constexpr int kNin = kMlpIn, kNh = kMlpHidden, kNout = kMlpOut;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kNin * kNh;
constexpr int kOffW2 = kOffB1 + kNh;
constexpr int kOffB2 = kOffW2 + kNh * kNh;
constexpr int kOffW3 = kOffB2 + kNh;
constexpr int kOffB3 = kOffW3 + kNh * kNout;
static_assert(kOffB3 + kNout == kMlpParams, "parameter layout");

// Row layout of one pair in the gradient staging buffer:
// [x(9), 1, h1(64), h2(64), dz1(64), dz2(64), dm(3)].
constexpr int kRowX = 0, kRowOne = 9, kRowH1 = 10, kRowH2 = kRowH1 + kNh, kRowDz1 = kRowH2 + kNh,
              kRowDz2 = kRowDz1 + kNh, kRowDm = kRowDz2 + kNh, kRowLen = kRowDm + kNout;
constexpr int kGradRows = 16;  // rows staged in LDS per sub-chunk

// MlpPackLayout: offsets of the per-lane MFMA operand images (bf16 for A,
// float32 for the bias-initialised accumulators C).
constexpr int kPackA1 = 0, kPackA2 = kPackA1 + 4 * 64 * 8, kPackA3 = kPackA2 + 4 * 2 * 64 * 8;
static_assert(kPackA3 + 2 * 64 * 8 == kMlpPackA, "weight pack layout");
constexpr int kPackC1 = 0, kPackC2 = kPackC1 + 4 * 64 * 4, kPackC3 = kPackC2 + 4 * 64 * 4;
static_assert(kPackC3 + 64 * 4 == kMlpPackC, "bias pack layout");

__device__ __forceinline__ uint32_t scene_of(const SceneDev* __restrict__ scenes, int nscenes, uint32_t s)
{
    int lo = 0, hi = nscenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s < scenes[mid].p_end)
            hi = mid;
        else
            lo = mid + 1;
    }
    return (uint32_t)lo;
}

// Features of pair (i = slot s, j) in float32 (the pressure term is the
// force_comp.glsl:59 factor (p_i+p_j)/(2 rho_j)).
// xj.w holds p_j and vj.w 1/rho_j, as the density pass leaves them.
__device__ __forceinline__ void pair_features(const SceneDev& sc, const MlpNorm& nm, float4 xi, float4 vi, float rhoi,
                                              float pi, float4 xj, float4 vj, float rhoj, float f[kNin])
{
    const float inv_h = 1.0f / sc.h;
    f[0] = (xi.x - xj.x) * inv_h;
    f[1] = (xi.y - xj.y) * inv_h;
    f[2] = (xi.z - xj.z) * inv_h;
    f[3] = (vi.x - vj.x) * nm.v_scale;
    f[4] = (vi.y - vj.y) * nm.v_scale;
    f[5] = (vi.z - vj.z) * nm.v_scale;
    f[6] = rhoi * nm.rho_scale;
    f[7] = rhoj * nm.rho_scale;
    f[8] = (pi + xj.w) * (0.5f * vj.w) * nm.pi_scale;
}

template <int SUB, bool EMIT>
__global__ __launch_bounds__(256) void k_pairs(const SceneDev* __restrict__ scenes, int nscenes,
                                               const uint32_t* __restrict__ start, const float4* __restrict__ P,
                                               const float4* __restrict__ V, const float* __restrict__ rho,
                                               uint32_t first, uint32_t stride,
                                               uint32_t T, MlpNorm nm, uint32_t* __restrict__ cnt,
                                               const uint64_t* __restrict__ off, uint32_t* __restrict__ pj,
                                               uint32_t* __restrict__ pt, float* __restrict__ X)
{
    __shared__ uint32_t rows[2 * (2 * SUB + 1) * 256];
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= T) return;
    const uint32_t s = first + t * stride;
    const SceneDev& sc = scenes[scene_of(scenes, nscenes, s)];
    const float4 xi = P[s];
    const float h2 = sc.h2;
    uint32_t count = 0;
    uint64_t row = EMIT ? off[t] : 0;
    float4 vi;
    float rhoi = 0.f, pi = 0.f;
    if constexpr (EMIT) {
        vi = V[s];
        rhoi = rho[s];
        pi = xi.w;
    }
    int cx, cy, cz;
    float fx, fy, fz;
    if (cell_coords(sc, xi.x, xi.y, xi.z, fx, fy, fz, cx, cy, cz)) {
        for_each_row<SUB, true, 256>(sc, start, fx, fy, fz, cx, cy, cz, rows + threadIdx.x, [&](uint32_t b, uint32_t e) {
            for (uint32_t j = b; j < e; ++j) {
                const float4 xj = P[j];
                const float ddx = xi.x - xj.x, ddy = xi.y - xj.y, ddz = xi.z - xj.z;
                const float r2 = fmaf(ddx, ddx, fmaf(ddy, ddy, ddz * ddz));
                if (r2 < h2 && j != s) {
                    if constexpr (EMIT) {
                        float f[kNin];
                        pair_features(sc, nm, xi, vi, rhoi, pi, xj, V[j], rho[j], f);
#pragma unroll
                        for (int k = 0; k < kNin; ++k) X[row * kNin + k] = f[k];
                        pj[row] = j;
                        pt[row] = t;
                        ++row;
                    } else {
                        ++count;
                    }
                }
            }
        });
    }
    if constexpr (!EMIT) cnt[t] = count;
}

__device__ __forceinline__ float gelu_poly(float x)
{
    const float s = fminf(x * x, kGeluClamp);
    const float q = fmaf(fmaf(fmaf(kGeluC3, s, kGeluC2), s, kGeluC1), s, kGeluC0) * s;
    return fmaf(0.5f, x, q) + 0.5f * fmaxf(fabsf(x) - kGeluKnee, 0.f);
}

__device__ __forceinline__ float gelu_poly_grad(float x)
{
    const float s = x * x;
    const float dq = fmaf(fmaf(fmaf(4.f * kGeluC3, s, 3.f * kGeluC2), s, 2.f * kGeluC1), s, kGeluC0);
    float g = 0.5f;
    if (s < kGeluClamp) g = fmaf(dq * 2.f, x, g);
    if (fabsf(x) > kGeluKnee) g += copysignf(0.5f, x);
    return g;
}

__device__ __forceinline__ void load_params_lds(float* lds, const float* __restrict__ params)
{
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) lds[i] = params[i];
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_mlp_forward_rows(const float* __restrict__ params, const float* __restrict__ X,
                                                          uint64_t P, float* __restrict__ Z1, float* __restrict__ Z2,
                                                          float* __restrict__ M)
{
    __shared__ float w[kMlpParams];
    load_params_lds(w, params);
    const uint64_t r = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (r >= P) return;
    float f[kNin];
#pragma unroll
    for (int k = 0; k < kNin; ++k) f[k] = X[r * kNin + k];
    float h1[kNh];
#pragma unroll
    for (int u = 0; u < kNh; ++u) {
        float z = w[kOffB1 + u];
#pragma unroll
        for (int k = 0; k < kNin; ++k) z = fmaf(f[k], w[kOffW1 + k * kNh + u], z);
        Z1[r * kNh + u] = z;
        h1[u] = gelu_poly(z);
    }
    float m[kNout] = {w[kOffB3], w[kOffB3 + 1], w[kOffB3 + 2]};
    for (int u = 0; u < kNh; ++u) {
        float z = w[kOffB2 + u];
#pragma unroll
        for (int k = 0; k < kNh; ++k) z = fmaf(h1[k], w[kOffW2 + k * kNh + u], z);
        Z2[r * kNh + u] = z;
        const float h = gelu_poly(z);
#pragma unroll
        for (int c = 0; c < kNout; ++c) m[c] = fmaf(h, w[kOffW3 + u * kNout + c], m[c]);
    }
#pragma unroll
    for (int c = 0; c < kNout; ++c) M[r * kNout + c] = m[c];
}

// pred_t = sum of the target's messages (pair order); loss term and dL/dpred.
__global__ __launch_bounds__(256) void k_mlp_aggregate(const float* __restrict__ M, const uint64_t* __restrict__ off,
                                                       uint32_t T, const float* __restrict__ tgt,
                                                       const float* __restrict__ wt, float inv_wsum,
                                                       float* __restrict__ pred, float* __restrict__ dpred,
                                                       double* __restrict__ lossp)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= T) return;
    float a[kNout] = {0.f, 0.f, 0.f};
    for (uint64_t r = off[t]; r < off[t + 1]; ++r)
#pragma unroll
        for (int c = 0; c < kNout; ++c) a[c] += M[r * kNout + c];
    double l = 0.0;
#pragma unroll
    for (int c = 0; c < kNout; ++c) {
        pred[t * kNout + c] = a[c];
        const float d = a[c] - tgt[t * kNout + c];
        dpred[t * kNout + c] = 2.f * wt[t] * d * inv_wsum;
        l += (double)wt[t] * (double)d * (double)d;
    }
    lossp[t] = l;
}

__global__ __launch_bounds__(256) void k_mlp_backward_rows(const float* __restrict__ params, const float* __restrict__ Z1,
                                                           const float* __restrict__ Z2, const uint32_t* __restrict__ pt,
                                                           const float* __restrict__ dpred, uint64_t P,
                                                           float* __restrict__ dZ1, float* __restrict__ dZ2)
{
    __shared__ float w[kMlpParams];
    load_params_lds(w, params);
    const uint64_t r = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (r >= P) return;
    const uint32_t t = pt[r];
    const float dm[kNout] = {dpred[t * kNout], dpred[t * kNout + 1], dpred[t * kNout + 2]};
    float dz2[kNh];
#pragma unroll
    for (int k = 0; k < kNh; ++k) {
        float d = 0.f;
#pragma unroll
        for (int c = 0; c < kNout; ++c) d = fmaf(w[kOffW3 + k * kNout + c], dm[c], d);
        dz2[k] = d * gelu_poly_grad(Z2[r * kNh + k]);
        dZ2[r * kNh + k] = dz2[k];
    }
    for (int k = 0; k < kNh; ++k) {
        float d = 0.f;
#pragma unroll
        for (int u = 0; u < kNh; ++u) d = fmaf(w[kOffW2 + k * kNh + u], dz2[u], d);
        dZ1[r * kNh + k] = d * gelu_poly_grad(Z1[r * kNh + k]);
    }
}

// Block b sums, over its rows [b*R, (b+1)*R), the outer products that make up
// every parameter's gradient; partial[b][p].
__global__ __launch_bounds__(256) void k_mlp_wgrad_partial(const float* __restrict__ X, const float* __restrict__ Z1,
                                                           const float* __restrict__ Z2, const float* __restrict__ dZ1,
                                                           const float* __restrict__ dZ2, const uint32_t* __restrict__ pt,
                                                           const float* __restrict__ dpred, uint64_t P, uint32_t R,
                                                           float* __restrict__ partial)
{
    constexpr int kOwn = (kMlpParams + 255) / 256;
    __shared__ float rowbuf[kGradRows * kRowLen];
    int ia[kOwn], ib[kOwn];
    float acc[kOwn];
#pragma unroll
    for (int q = 0; q < kOwn; ++q) {
        const int p = threadIdx.x + 256 * q;
        acc[q] = 0.f;
        ia[q] = kRowOne;
        ib[q] = kRowOne;
        if (p < kOffB1) {
            ia[q] = kRowX + p / kNh;
            ib[q] = kRowDz1 + p % kNh;
        } else if (p < kOffW2) {
            ib[q] = kRowDz1 + (p - kOffB1);
        } else if (p < kOffB2) {
            ia[q] = kRowH1 + (p - kOffW2) / kNh;
            ib[q] = kRowDz2 + (p - kOffW2) % kNh;
        } else if (p < kOffW3) {
            ib[q] = kRowDz2 + (p - kOffB2);
        } else if (p < kOffB3) {
            ia[q] = kRowH2 + (p - kOffW3) / kNout;
            ib[q] = kRowDm + (p - kOffW3) % kNout;
        } else if (p < kMlpParams) {
            ib[q] = kRowDm + (p - kOffB3);
        }
    }
    const uint64_t r0 = (uint64_t)blockIdx.x * R;
    const uint64_t r1 = r0 + R < P ? r0 + R : P;
    for (uint64_t rb = r0; rb < r1; rb += kGradRows) {
        const int nr = (int)((r1 - rb) < (uint64_t)kGradRows ? (r1 - rb) : (uint64_t)kGradRows);
        __syncthreads();
        for (int i = threadIdx.x; i < kGradRows * kRowLen; i += 256) {
            const int rr = i / kRowLen, c = i % kRowLen;
            float v = 0.f;
            if (rr < nr) {
                const uint64_t r = rb + rr;
                if (c < kRowOne)
                    v = X[r * kNin + c];
                else if (c == kRowOne)
                    v = 1.f;
                else if (c < kRowH2)
                    v = gelu_poly(Z1[r * kNh + (c - kRowH1)]);
                else if (c < kRowDz1)
                    v = gelu_poly(Z2[r * kNh + (c - kRowH2)]);
                else if (c < kRowDz2)
                    v = dZ1[r * kNh + (c - kRowDz1)];
                else if (c < kRowDm)
                    v = dZ2[r * kNh + (c - kRowDz2)];
                else
                    v = dpred[pt[r] * kNout + (c - kRowDm)];
            }
            rowbuf[i] = v;
        }
        __syncthreads();
        for (int rr = 0; rr < nr; ++rr) {
            const float* row = rowbuf + rr * kRowLen;
#pragma unroll
            for (int q = 0; q < kOwn; ++q) acc[q] = fmaf(row[ia[q]], row[ib[q]], acc[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < kOwn; ++q) {
        const int p = threadIdx.x + 256 * q;
        if (p < kMlpParams) partial[(size_t)blockIdx.x * kMlpParams + p] = acc[q];
    }
}

__global__ __launch_bounds__(256) void k_mlp_wgrad_reduce(const float* __restrict__ partial, uint32_t nblocks,
                                                          float* __restrict__ grad)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kMlpParams) return;
    double s = 0.0;
    for (uint32_t b = 0; b < nblocks; ++b) s += (double)partial[(size_t)b * kMlpParams + p];
    grad[p] = (float)s;
}

// Adam (Kingma and Ba, 2015) with bias correction, float32 master weights.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ params, const float* __restrict__ grad,
                                                  float* __restrict__ m, float* __restrict__ v, float lr, float beta1,
                                                  float beta2, float eps, float bc1, float bc2)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kMlpParams) return;
    const float g = grad[p];
    const float mn = beta1 * m[p] + (1.f - beta1) * g;
    const float vn = beta2 * v[p] + (1.f - beta2) * g * g;
    m[p] = mn;
    v[p] = vn;
    params[p] -= lr * (mn / bc1) / (sqrtf(vn / bc2) + eps);
}

// Training targets of target t (slot first + t*stride): the mean acceleration
// over the next 4 fine steps, (v(t+4dt) - v(t))/(4 dt) - G, in units of
// acc_scale; weight 0 for particles that are non-finite or within 1e-6 of a
// wall in either state (the wall clamp is not part of the learned update).
__global__ __launch_bounds__(256) void k_train_targets(const SceneDev* __restrict__ scenes, int nscenes,
                                                       const uint32_t* __restrict__ perm,
                                                       const sph_particle_t* __restrict__ now,
                                                       const sph_particle_t* __restrict__ later, uint32_t first,
                                                       uint32_t stride, uint32_t T, float inv_acc_scale,
                                                       float* __restrict__ tgt, float* __restrict__ wt)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= T) return;
    const uint32_t s = first + t * stride;
    const SceneDev& sc = scenes[scene_of(scenes, nscenes, s)];
    const uint32_t o = perm[s];
    const sph_particle_t& a = now[o];
    const sph_particle_t& b = later[o];
    bool ok = true;
    float out[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float acc = (b.vel[c] - a.vel[c]) / (4.f * sc.dt) - sc.grav[c];
        out[c] = acc * inv_acc_scale;
        ok = ok && isfinite(out[c]) && isfinite(a.pos[c]) && isfinite(b.pos[c]);
        const float p0 = a.pos[c], p1 = b.pos[c];
        ok = ok && fabsf(p0 - sc.lower[c]) > 1e-6f && fabsf(p0 - sc.upper[c]) > 1e-6f &&
             fabsf(p1 - sc.lower[c]) > 1e-6f && fabsf(p1 - sc.upper[c]) > 1e-6f;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) tgt[t * 3 + c] = ok ? out[c] : 0.f;
    wt[t] = ok ? 1.f : 0.f;
}

// Coarse step for every slot: a = acc_scale*pred + G (network) or a = F/rho
// (SPH force held over the 4 fine steps); v' = v + 4dt a,
// x' = x + 4dt v + 10 dt^2 a, then the wall clamp of integrate_comp.glsl:46-77.
// Writes the AoS record (pos, vel, force = rho*a, extras.xy) and state B.
__global__ __launch_bounds__(kTile) void k_coarse_integrate(const Tile* __restrict__ tiles,
                                                            const SceneDev* __restrict__ scenes,
                                                            const float4* __restrict__ PA, const float4* __restrict__ VA,
                                                            const float* __restrict__ rho,
                                                            const float* __restrict__ pred,
                                                            const uint32_t* __restrict__ perm,
                                                            sph_particle_t* __restrict__ aos, float4* __restrict__ PB,
                                                            float4* __restrict__ VB, float acc_scale, int use_force)
{
#pragma clang fp contract(off)
    const Tile t = tiles[blockIdx.x];
    const SceneDev& sc = scenes[t.scene];
    const uint32_t s = t.begin + threadIdx.x;
    if (s >= t.end) return;
    const float4 x = PA[s], v = VA[s];
    const float rh = rho[s];
    float* r = reinterpret_cast<float*>(aos + perm[s]);
    float a[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) a[c] = use_force ? r[8 + c] / rh : acc_scale * pred[s * 3 + c] + sc.grav[c];
    const float xs[3] = {x.x, x.y, x.z}, vs[3] = {v.x, v.y, v.z};
    const float dt4 = 4.f * sc.dt, dt2_10 = 10.f * sc.dt * sc.dt;
    float nx[3], nv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float vel = vs[c] + dt4 * a[c];
        float pos = xs[c] + dt4 * vs[c] + dt2_10 * a[c];
        if (pos < sc.lower[c]) {
            pos = sc.lower[c];
            vel *= -sc.damping;
        } else if (pos > sc.upper[c]) {
            pos = sc.upper[c];
            vel *= -sc.damping;
        }
        nx[c] = pos;
        nv[c] = vel;
    }
    PB[s] = make_float4(nx[0], nx[1], nx[2], 0.f);
    VB[s] = make_float4(nv[0], nv[1], nv[2], 0.f);
    r[0] = nx[0];
    r[1] = nx[1];
    r[2] = nx[2];
    r[4] = nv[0];
    r[5] = nv[1];
    r[6] = nv[2];
    r[8] = rh * a[0];
    r[9] = rh * a[1];
    r[10] = rh * a[2];
    r[12] = rh;
    r[13] = x.w;
}

// Summed messages of every slot, FP32, one particle per lane and the whole MLP
// per pair in registers (weights in LDS).  The float32 reference of the
// coarse step; the fast path is the MFMA kernel.
template <int SUB>
__global__ __launch_bounds__(256) void k_coarse_mlp_fp32(const Tile* __restrict__ tiles,
                                                         const SceneDev* __restrict__ scenes,
                                                         const uint32_t* __restrict__ start,
                                                         const float4* __restrict__ P, const float4* __restrict__ V,
                                                         const float* __restrict__ rho,
                                                         const float* __restrict__ params, MlpNorm nm,
                                                         float* __restrict__ pred)
{
    __shared__ float w[kMlpParams];
    __shared__ uint32_t rows[2 * (2 * SUB + 1) * 256];
    load_params_lds(w, params);
    const Tile t = tiles[blockIdx.x];
    const SceneDev& sc = scenes[t.scene];
    const uint32_t s = t.begin + threadIdx.x;
    if (s >= t.end) return;
    const float4 xi = P[s], vi = V[s];
    const float rhoi = rho[s], pi = xi.w;
    const float h2 = sc.h2;
    float acc[kNout] = {0.f, 0.f, 0.f};
    int cx, cy, cz;
    float fx, fy, fz;
    if (cell_coords(sc, xi.x, xi.y, xi.z, fx, fy, fz, cx, cy, cz)) {
        for_each_row<SUB, true, 256>(sc, start, fx, fy, fz, cx, cy, cz, rows + threadIdx.x, [&](uint32_t b, uint32_t e) {
            for (uint32_t j = b; j < e; ++j) {
                const float4 xj = P[j];
                const float ddx = xi.x - xj.x, ddy = xi.y - xj.y, ddz = xi.z - xj.z;
                const float r2 = fmaf(ddx, ddx, fmaf(ddy, ddy, ddz * ddz));
                if (!(r2 < h2) || j == s) continue;
                float f[kNin];
                pair_features(sc, nm, xi, vi, rhoi, pi, xj, V[j], rho[j], f);
                float h1[kNh];
#pragma unroll
                for (int u = 0; u < kNh; ++u) {
                    float z = w[kOffB1 + u];
#pragma unroll
                    for (int k = 0; k < kNin; ++k) z = fmaf(f[k], w[kOffW1 + k * kNh + u], z);
                    h1[u] = gelu_poly(z);
                }
                float m[kNout] = {w[kOffB3], w[kOffB3 + 1], w[kOffB3 + 2]};
                for (int u = 0; u < kNh; ++u) {
                    float z = w[kOffB2 + u];
#pragma unroll
                    for (int k = 0; k < kNh; ++k) z = fmaf(h1[k], w[kOffW2 + k * kNh + u], z);
                    const float hh = gelu_poly(z);
#pragma unroll
                    for (int c = 0; c < kNout; ++c) m[c] = fmaf(hh, w[kOffW3 + u * kNout + c], m[c]);
                }
#pragma unroll
                for (int c = 0; c < kNout; ++c) acc[c] += m[c];
            }
        });
    }
#pragma unroll
    for (int c = 0; c < kNout; ++c) pred[s * 3 + c] = acc[c];
}

// ---- BF16 MFMA inference --------------------------------------------------
//
// v_mfma_f32_16x16x32_bf16 lane maps (lane l, element e):
//   A (16 rows x 32 k):  row l%16, k 8*(l/16)+e
//   B (32 k x 16 cols):  col l%16, k 8*(l/16)+e
//   C/D (16 x 16 fp32):  row 4*(l/16)+e, col l%16
// The network runs transposed, hidden units x pairs: H1' = W1' X' + b1,
// H2' = W2' g(H1') + b2, M' = W3' g(H2') + b3, 16 pairs per tile.  A layer's
// D registers hold, per lane, pair l%16 and hidden units 16m + 4(l/16) + e
// (m = 0..3 tiles).  Those are exactly the B elements of k-step s of the next
// layer if its K order is permuted as
//   k-slot 32s + 8g + 4mm + e  <->  hidden unit 16(2s+mm) + 4g + e,
// so layers chain in registers; the permutation is folded into the packed
// weights (k_mlp_pack).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// GELU of the MFMA kernel by table lookup.  An activation is converted to FP16
// (round toward zero) and the top kGeluBits bits of its FP16 pattern (sign, 5
// exponent bits, 6 mantissa bits) index the table; entry i holds gelu_poly at
// the centre of that bucket, rounded to BF16 (the next MFMA's operand type).
// Over the whole FP16 range this stays within about one BF16 step of
// gelu_poly, the training-time form, at one conversion, two shifts and two
// LDS reads per pair of activations.  The 4,096-entry table (8 KB of LDS)
// keeps 6 workgroups of 4 waves per CU; a 13-bit table would halve that.
constexpr int kGeluBits = 12;
constexpr int kGeluTab = 1 << kGeluBits;
constexpr int kGeluShift = 16 - kGeluBits;
static_assert(kGeluTab <= kMlpGeluTab, "table buffer");

__global__ __launch_bounds__(256) void k_gelu_table(uint16_t* __restrict__ tab)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= (uint32_t)kGeluTab) return;
    const uint32_t bits = (i << kGeluShift) | (1u << (kGeluShift - 1));  // bucket centre
    const uint32_t expo = (bits >> 10) & 31u, sign = bits >> 15;
    float g;
    if (expo == 31u) {
        // Infinities (and the NaNs sharing their bucket): GELU(+inf) = +inf,
        // GELU(-inf) = 0; the other buckets of this exponent are NaN.
        const bool inf_bucket = ((bits & 0x3ffu) >> kGeluShift) == 0u;
        g = inf_bucket ? (sign ? 0.f : __builtin_inff()) : __builtin_nanf("");
    } else {
        g = gelu_poly((float)__builtin_bit_cast(_Float16, (uint16_t)bits));
    }
    tab[i] = __builtin_bit_cast(uint16_t, (__bf16)g);
}

__device__ __forceinline__ bf16x2 gelu2_tab(const uint16_t* tab, float a, float b)
{
    const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(a, b));
    const uint32_t lo = tab[(h & 0xffffu) >> kGeluShift], hi = tab[h >> (16 + kGeluShift)];
    return __builtin_bit_cast(bf16x2, lo | (hi << 16));
}

__host__ __device__ __forceinline__ int kslot_unit(int kslot)
{
    const int s = kslot / 32, g = (kslot % 32) / 8, mm = (kslot % 8) / 4, e = kslot % 4;
    return 16 * (2 * s + mm) + 4 * g + e;
}

// Per-lane operand images of the network (see MlpPackLayout).
__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ params, __bf16* __restrict__ pa,
                                                  float* __restrict__ pc)
{
    // A1[m][l][e]: W1[k][16m + l%16], k = 8(l/16)+e < 9
    for (int i = threadIdx.x; i < 4 * 64 * 8; i += 256) {
        const int m = i / 512, l = (i / 8) % 64, e = i % 8, k = 8 * (l / 16) + e;
        pa[kPackA1 + i] = (__bf16)(k < kNin ? params[kOffW1 + k * kNh + 16 * m + l % 16] : 0.f);
    }
    // A2[m][s][l][e]: W2[unit(32s + 8(l/16) + e)][16m + l%16]
    for (int i = threadIdx.x; i < 4 * 2 * 64 * 8; i += 256) {
        const int m = i / 1024, s = (i / 512) % 2, l = (i / 8) % 64, e = i % 8;
        const int u = kslot_unit(32 * s + 8 * (l / 16) + e);
        pa[kPackA2 + i] = (__bf16)params[kOffW2 + u * kNh + 16 * m + l % 16];
    }
    // A3[s][l][e]: W3[unit(32s + 8(l/16) + e)][l%16] for l%16 < 3
    for (int i = threadIdx.x; i < 2 * 64 * 8; i += 256) {
        const int s = i / 512, l = (i / 8) % 64, e = i % 8;
        const int u = kslot_unit(32 * s + 8 * (l / 16) + e);
        pa[kPackA3 + i] = (__bf16)(l % 16 < kNout ? params[kOffW3 + u * kNout + l % 16] : 0.f);
    }
    // C1[m][l][e] = b1[16m + 4(l/16) + e], C2 likewise with b2, C3[l][e] = b3[4(l/16)+e] (rows < 3)
    for (int i = threadIdx.x; i < 4 * 64 * 4; i += 256) {
        const int m = i / 256, l = (i / 4) % 64, e = i % 4, u = 16 * m + 4 * (l / 16) + e;
        pc[kPackC1 + i] = params[kOffB1 + u];
        pc[kPackC2 + i] = params[kOffB2 + u];
    }
    for (int i = threadIdx.x; i < 64 * 4; i += 256) {
        const int l = i / 4, e = i % 4, row = 4 * (l / 16) + e;
        pc[kPackC3 + i] = row < kNout ? params[kOffB3 + row] : 0.f;
    }
}

constexpr int kHitCap = 192;  // hit list entries per wave

// Summed messages of every slot on MFMA: one wave per target particle.  The
// wave scans the target's stencil rows 64 candidates at a time, compacts the
// hits into an LDS list, and runs the network on 16-pair tiles.  The network's
// operands, the hit lists and the GELU table take 26 KB of LDS per 4-wave
// workgroup, so 6 waves share each SIMD: the GELU lookups of one wave issue
// while others wait on their MFMA results and loads.  The register budget of
// 6 waves (85 per lane) lets the MFMA results land in VGPRs, where the GELU
// reads them; at a budget of 64 they went through AGPRs and cost 35 extra
// moves per 16-pair tile.
template <int SUB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_coarse_mlp_mfma(const SceneDev* __restrict__ scenes, int nscenes,
                                                         const uint32_t* __restrict__ start,
                                                         const float4* __restrict__ P, const float4* __restrict__ V,
                                                         const float* __restrict__ rho,
                                                         const __bf16* __restrict__ pa, const float* __restrict__ pc,
                                                         MlpNorm nm, uint32_t n, float* __restrict__ pred,
                                                         unsigned long long* __restrict__ pair_count,
                                                         const uint16_t* __restrict__ gtab)
{
    __shared__ uint32_t hits_all[4][kHitCap];
    // The GELU table (k_gelu_table), staged once per workgroup.
    __shared__ uint4 s_gelu4[kGeluTab / 8];
    const uint16_t* s_gelu = reinterpret_cast<const uint16_t*>(s_gelu4);
    for (int i = (int)threadIdx.x; i < kGeluTab / 8; i += 256) s_gelu4[i] = reinterpret_cast<const uint4*>(gtab)[i];
    // The network's MFMA operands (weights and biases), staged once per
    // workgroup as per-lane images and read back at each MFMA, so that a
    // wave's registers hold only its working set (more waves per SIMD).
    __shared__ bf16x8 s_a1[4][64], s_a2[8][64], s_a3[2][64];
    // Biases once each: b1[64], b2[64], b3 padded with zeros to 16.  A lane's
    // C operand for MFMA m is the float4 at 16m + 4(lane/16) (a 16-lane
    // broadcast), the same values as its packed image.
    __shared__ f32x4 s_b[16 + 16 + 4];
    uint32_t wave_pairs = 0;  // pairs this wave evaluated, added to *pair_count at the end
    const int lane = (int)(threadIdx.x & 63);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    uint32_t* hits = hits_all[wv];
    for (int i = (int)threadIdx.x; i < 4 * 64; i += 256) {
        const int m = i / 64, l = i % 64;
        bf16x8 a;
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] = pa[kPackA1 + (m * 64 + l) * 8 + e];
        s_a1[m][l] = a;
    }
    if (threadIdx.x < 36) {
        // float4 q of s_b: bias rows 4q..4q+3 of b1 (q < 16), b2 (q < 32) or b3,
        // read from the packed image of lane 16*(q%4) of MFMA q/4.
        const int q = (int)threadIdx.x, base = q < 16 ? kPackC1 : (q < 32 ? kPackC2 : kPackC3);
        const int r = q < 32 ? q % 16 : q - 32;
        const int img = q < 32 ? ((r / 4) * 64 + 16 * (r % 4)) * 4 : 16 * r * 4;
        f32x4 c;
#pragma unroll
        for (int e = 0; e < 4; ++e) c[e] = pc[base + img + e];
        s_b[q] = c;
    }
    for (int i = (int)threadIdx.x; i < 2 * 64; i += 256) {
        const int m = i / 64, l = i % 64;
        bf16x8 a;
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] = pa[kPackA3 + (m * 64 + l) * 8 + e];
        s_a3[m][l] = a;
    }
    for (int i = (int)threadIdx.x; i < 8 * 64; i += 256) {
        const int ms = i / 64, l = i % 64;
        bf16x8 a;
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] = pa[kPackA2 + (ms * 64 + l) * 8 + e];
        s_a2[ms][l] = a;
    }
    __syncthreads();

    const uint32_t nwaves = gridDim.x * 4u;
    for (uint32_t si = blockIdx.x * 4u + (uint32_t)wv; si < n; si += nwaves) {
        const uint32_t s = (uint32_t)__builtin_amdgcn_readfirstlane((int)si);
        const SceneDev& sc = scenes[scene_of(scenes, nscenes, s)];
        const float4 xi = P[s], vi = V[s];
        const float rhoi = rho[s], pi = xi.w;
        const float h2 = sc.h2, inv_h = 1.0f / sc.h;
        float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
        int nh = 0;

        // Run the network on the first `count` listed hits in 16-pair tiles
        // (all lanes take part); with `partial`, a last incomplete tile too.
        auto run_tiles = [&](int count, bool partial) {
            const int ntiles = partial ? (count + 15) / 16 : count / 16;
            for (int tt = 0; tt < ntiles; ++tt) {
                // Keeps the staged operands' LDS reads inside the loop (no
                // instruction is emitted).
                asm volatile("" ::: "memory");
                const int pidx = tt * 16 + (lane & 15);
                const bool valid = pidx < count;
                const uint32_t j = valid ? hits[pidx] : s;
                bf16x8 bx;
#pragma unroll
                for (int e = 0; e < 8; ++e) bx[e] = (__bf16)0.f;
                if (lane < 16) {
                    const float4 xj = P[j], vj = V[j];
                    const float rj = rho[j];
                    bx[0] = (__bf16)((xi.x - xj.x) * inv_h);
                    bx[1] = (__bf16)((xi.y - xj.y) * inv_h);
                    bx[2] = (__bf16)((xi.z - xj.z) * inv_h);
                    bx[3] = (__bf16)((vi.x - vj.x) * nm.v_scale);
                    bx[4] = (__bf16)((vi.y - vj.y) * nm.v_scale);
                    bx[5] = (__bf16)((vi.z - vj.z) * nm.v_scale);
                    bx[6] = (__bf16)(rhoi * nm.rho_scale);
                    bx[7] = (__bf16)(rj * nm.rho_scale);
                } else if (lane < 32) {
                    const float pj = P[j].w, inv_rhoj = V[j].w;
                    bx[0] = (__bf16)((pi + pj) * (0.5f * inv_rhoj) * nm.pi_scale);
                }
                f32x4 h1[4];
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    h1[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(s_a1[m][lane], bx, s_b[4 * m + lane / 16], 0, 0, 0);
                bf16x8 b2[2];
#pragma unroll
                for (int ks = 0; ks < 2; ++ks)
#pragma unroll
                    for (int mm = 0; mm < 2; ++mm)
#pragma unroll
                        for (int e = 0; e < 4; e += 2) {
                            const bf16x2 g = gelu2_tab(s_gelu, h1[2 * ks + mm][e], h1[2 * ks + mm][e + 1]);
                            b2[ks][4 * mm + e] = g.x;
                            b2[ks][4 * mm + e + 1] = g.y;
                        }
                f32x4 h2[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    h2[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(s_a2[2 * m][lane], b2[0], s_b[16 + 4 * m + lane / 16], 0, 0, 0);
                    h2[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(s_a2[2 * m + 1][lane], b2[1], h2[m], 0, 0, 0);
                }
                bf16x8 b3[2];
#pragma unroll
                for (int ks = 0; ks < 2; ++ks)
#pragma unroll
                    for (int mm = 0; mm < 2; ++mm)
#pragma unroll
                        for (int e = 0; e < 4; e += 2) {
                            const bf16x2 g = gelu2_tab(s_gelu, h2[2 * ks + mm][e], h2[2 * ks + mm][e + 1]);
                            b3[ks][4 * mm + e] = g.x;
                            b3[ks][4 * mm + e + 1] = g.y;
                        }
                f32x4 o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(s_a3[0][lane], b3[0], s_b[32 + lane / 16], 0, 0, 0);
                o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(s_a3[1][lane], b3[1], o, 0, 0, 0);
                if (lane < 16 && valid) {
                    acc0 += o[0];
                    acc1 += o[1];
                    acc2 += o[2];
                }
            }
            // keep the unprocessed tail at the front of the list
            const int done = ntiles * 16;
            const int rest = count - done;
            if (!partial && rest > 0) {
                const uint32_t v = lane < rest ? hits[done + lane] : 0u;
                if (lane < rest) hits[lane] = v;
            }
            return partial ? 0 : rest;
        };

        int cx, cy, cz;
        float fx, fy, fz;
        if (cell_coords(sc, xi.x, xi.y, xi.z, fx, fy, fz, cx, cy, cz)) {
            const int dx = sc.dims[0], dy = sc.dims[1], dz = sc.dims[2];
            constexpr int SUBX = kXSplit * SUB;
            const int xa = cx > SUBX ? cx - SUBX : 0, xb = cx < dx - SUBX ? cx + SUBX : dx - 1;
            const int z0 = cz > SUB ? cz - SUB : 0, z1 = cz < dz - SUB ? cz + SUB : dz - 1;
            const int y0 = cy > SUB ? cy - SUB : 0, y1 = cy < dy - SUB ? cy + SUB : dy - 1;
            constexpr float kReach2 = (float)(SUB * SUB);
            // Lane r < nrows finds stencil row r (z-major, y ascending, as the
            // plain walk orders them): its slot range [rb, rb + rl), empty when
            // the row lies beyond h.  At most (2 SUB + 1)^2 <= 49 rows.
            static_assert((2 * SUB + 1) * (2 * SUB + 1) <= 64, "one lane per stencil row");
            const int ny = y1 - y0 + 1, nrows = (z1 - z0 + 1) * ny;
            uint32_t rb = 0, rl = 0;
            if (lane < nrows) {
                const int z = z0 + lane / ny, y = y0 + lane % ny;
                const float gz = z > cz ? (float)z - fz : (z < cz ? fz - (float)(z + 1) : 0.f);
                const float gy = y > cy ? (float)y - fy : (y < cy ? fy - (float)(y + 1) : 0.f);
                const float rem = kReach2 - fmaf(gy, gy, gz * gz);
                if (rem > 0.f) {
                    const float w = (float)kXSplit * __builtin_amdgcn_sqrtf(rem);
                    const int x0 = (int)fmaxf(fminf(floorf(fx - w), (float)(dx - 1)), (float)xa);
                    const int x1 = (int)fminf(fmaxf(floorf(fx + w), 0.f), (float)xb);
                    const uint32_t row = row_key(sc, y, z);
                    rb = start[row + (uint32_t)x0];
                    rl = start[row + (uint32_t)(x1 + 1)] - rb;
                }
            }
            // The rows laid end to end: row r covers positions [incl - rl, incl).
            uint32_t incl = rl;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t t = __shfl_up(incl, d);
                if (lane >= d) incl += t;
            }
            const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            // 64 positions per round, so rounds are full across row ends; the
            // hits keep the row-by-row, slot-ascending order.
            int r0 = 0;  // first row that ends after the round's first position
            for (uint32_t base = 0; base < total; base += 64u) {
                while ((uint32_t)__builtin_amdgcn_readlane((int)incl, r0) <= base) ++r0;
                const uint32_t g = base + (uint32_t)lane;
                uint32_t j = 0;
                for (int q = r0; q < nrows; ++q) {
                    const uint32_t qe = (uint32_t)__builtin_amdgcn_readlane((int)incl, q);
                    const uint32_t qs = qe - (uint32_t)__builtin_amdgcn_readlane((int)rl, q);
                    if (qs >= base + 64u) break;
                    if (g >= qs && g < qe) j = (uint32_t)__builtin_amdgcn_readlane((int)rb, q) + (g - qs);
                }
                bool hit = false;
                if (g < total && j != s) {
                    const float4 xj = P[j];
                    const float ddx = xi.x - xj.x, ddy = xi.y - xj.y, ddz = xi.z - xj.z;
                    hit = fmaf(ddx, ddx, fmaf(ddy, ddy, ddz * ddz)) < h2;
                }
                const uint64_t mask = __ballot(hit);
                const int posn = nh + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                if (hit) hits[posn] = j;
                nh += __popcll(mask);
                wave_pairs += (uint32_t)__popcll(mask);
                if (nh > kHitCap - 64) nh = run_tiles(nh, false);
            }
        }
        run_tiles(nh, true);
        // sum the 16 pair lanes (fixed tree order)
#pragma unroll
        for (int d = 8; d > 0; d >>= 1) {
            acc0 += __shfl_xor(acc0, d);
            acc1 += __shfl_xor(acc1, d);
            acc2 += __shfl_xor(acc2, d);
        }
        if (lane == 0) {
            pred[s * 3] = acc0;
            pred[s * 3 + 1] = acc1;
            pred[s * 3 + 2] = acc2;
        }
    }
    if (lane == 0 && wave_pairs != 0u) atomicAdd(pair_count, (unsigned long long)wave_pairs);
}

// pred of slot s scattered to AoS order (diagnostics).
__global__ __launch_bounds__(256) void k_pred_to_aos(const float* __restrict__ pred, const uint32_t* __restrict__ perm,
                                                     uint32_t n, float* __restrict__ out)
{
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = perm[s];
#pragma unroll
    for (int c = 0; c < 3; ++c) out[o * 3 + c] = pred[s * 3 + c];
}

inline unsigned blocks256(uint64_t n) { return (unsigned)((n + 255u) / 256u); }

}  // namespace

hipError_t launch_pairs(hipStream_t st, const DeviceState& d, int nscenes, const MlpWork& w, uint32_t first,
                        uint32_t stride, uint32_t T, bool emit)
{
    if (T == 0) return hipSuccess;
    const dim3 grid(blocks256(T)), block(256);
    const int sub = d.tuning.cells_per_h;
#define NPRSPH_PAIRS(S)                                                                                          \
    if (emit)                                                                                                    \
        hipLaunchKernelGGL((k_pairs<S, true>), grid, block, 0, st, d.scenes, nscenes, d.start, d.PA, d.VA, d.rho, \
                           first, stride, T, w.norm, w.cnt, w.off, w.pj, w.pt, w.X);                       \
    else                                                                                                         \
        hipLaunchKernelGGL((k_pairs<S, false>), grid, block, 0, st, d.scenes, nscenes, d.start, d.PA, d.VA, d.rho, \
                           first, stride, T, w.norm, w.cnt, w.off, w.pj, w.pt, w.X);
    if (sub == 1) {
        NPRSPH_PAIRS(1)
    } else if (sub == 2) {
        NPRSPH_PAIRS(2)
    } else {
        NPRSPH_PAIRS(3)
    }
#undef NPRSPH_PAIRS
    return hipGetLastError();
}

hipError_t launch_mlp_forward(hipStream_t st, const MlpWork& w, uint64_t P)
{
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_forward_rows, dim3(blocks256(P)), dim3(256), 0, st, w.params, w.X, P, w.Z1, w.Z2, w.M);
    return hipGetLastError();
}

hipError_t launch_mlp_aggregate(hipStream_t st, const MlpWork& w, uint32_t T, float inv_wsum)
{
    if (T == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_aggregate, dim3(blocks256(T)), dim3(256), 0, st, w.M, w.off, T, w.tgt, w.wt, inv_wsum,
                       w.pred, w.dpred, w.lossp);
    return hipGetLastError();
}

hipError_t launch_mlp_backward(hipStream_t st, const MlpWork& w, uint64_t P, uint32_t rows_per_block,
                               uint32_t nblocks)
{
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_backward_rows, dim3(blocks256(P)), dim3(256), 0, st, w.params, w.Z1, w.Z2, w.pt, w.dpred,
                       P, w.dZ1, w.dZ2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_wgrad_partial, dim3(nblocks), dim3(256), 0, st, w.X, w.Z1, w.Z2, w.dZ1, w.dZ2, w.pt,
                       w.dpred, P, rows_per_block, w.partial);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_wgrad_reduce, dim3(blocks256(kMlpParams)), dim3(256), 0, st, w.partial, nblocks, w.grad);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(hipStream_t st, const MlpWork& w, float lr, float beta1, float beta2, float eps, float bc1,
                           float bc2)
{
    hipLaunchKernelGGL(k_mlp_adam, dim3(blocks256(kMlpParams)), dim3(256), 0, st, w.params, w.grad, w.adam_m, w.adam_v,
                       lr, beta1, beta2, eps, bc1, bc2);
    return hipGetLastError();
}

hipError_t launch_train_targets(hipStream_t st, const DeviceState& d, int nscenes, const MlpWork& w,
                                const sph_particle_t* later, uint32_t first, uint32_t stride, uint32_t T)
{
    if (T == 0) return hipSuccess;
    hipLaunchKernelGGL(k_train_targets, dim3(blocks256(T)), dim3(256), 0, st, d.scenes, nscenes, d.perm, d.aos, later,
                       first, stride, T, 1.0f / w.norm.acc_scale, w.tgt, w.wt);
    return hipGetLastError();
}

hipError_t launch_coarse_integrate(hipStream_t st, const DeviceState& d, const MlpWork& w, bool use_force)
{
    hipLaunchKernelGGL(k_coarse_integrate, dim3(d.ntiles), dim3(kTile), 0, st, d.tiles, d.scenes, d.PA, d.VA, d.rho,
                       w.pred, d.perm, d.aos, d.PB, d.VB, w.norm.acc_scale, use_force ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_mlp_pack(hipStream_t st, const MlpWork& w)
{
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(256), 0, st, w.params, w.pack_a, w.pack_c);
    hipLaunchKernelGGL(k_gelu_table, dim3((kGeluTab + 255) / 256), dim3(256), 0, st, w.gelu_tab);
    return hipGetLastError();
}

hipError_t launch_coarse_mlp(hipStream_t st, const DeviceState& d, int nscenes, const MlpWork& w, int path)
{
    const int sub = d.tuning.cells_per_h;
    if (path == 1) {
        const hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(w.pairs), 2u, 0u);
        if (e != hipSuccess) return e;
        const dim3 grid(std::min<uint32_t>((d.n + 3u) / 4u, 256u * 64u)), block(256);
#define NPRSPH_MFMA(S)                                                                                             \
    hipLaunchKernelGGL(k_coarse_mlp_mfma<S>, grid, block, 0, st, d.scenes, nscenes, d.start, d.PA, d.VA, d.rho, \
                       w.pack_a, w.pack_c, w.norm, d.n, w.pred, w.pairs, w.gelu_tab)
        if (sub == 1)
            NPRSPH_MFMA(1);
        else if (sub == 2)
            NPRSPH_MFMA(2);
        else
            NPRSPH_MFMA(3);
#undef NPRSPH_MFMA
        return hipGetLastError();
    }
    if (sub == 1)
        hipLaunchKernelGGL(k_coarse_mlp_fp32<1>, dim3(d.ntiles), dim3(256), 0, st, d.tiles, d.scenes, d.start, d.PA,
                           d.VA, d.rho, w.params, w.norm, w.pred);
    else if (sub == 2)
        hipLaunchKernelGGL(k_coarse_mlp_fp32<2>, dim3(d.ntiles), dim3(256), 0, st, d.tiles, d.scenes, d.start, d.PA,
                           d.VA, d.rho, w.params, w.norm, w.pred);
    else
        hipLaunchKernelGGL(k_coarse_mlp_fp32<3>, dim3(d.ntiles), dim3(256), 0, st, d.tiles, d.scenes, d.start, d.PA,
                           d.VA, d.rho, w.params, w.norm, w.pred);
    return hipGetLastError();
}

hipError_t launch_pred_to_aos(hipStream_t st, const DeviceState& d, const MlpWork& w, float* out)
{
    hipLaunchKernelGGL(k_pred_to_aos, dim3(blocks256(d.n)), dim3(256), 0, st, w.pred, d.perm, d.n, out);
    return hipGetLastError();
}

}  // namespace nprsph
