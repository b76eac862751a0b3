// ff_learned.hip -- learned per-pixel fusion confidence on gfx950.
//
// The reference weighs every depth sample by a fixed linear fall-off
// (getDistanceWeight_AoS_variables_h, inline_functions.hpp:83-124; in SIMD
// form update_functions.cpp:1355-1403).  Here an optional per-pixel
// confidence c in (0, 1) multiplies that weight: wI' = wI * c(pixel).
// c = sigmoid(MLP(features)) with
//   features  25: 5x5 depth patch around the pixel, z_k / z_c - 1 clamped to
//                 [-1, 1] (missing or outside the image: -1)
//              1: z_c / 4 m
//              1: |cos| of the angle between the viewing ray and the normal
//                 from central differences (0 when a neighbour is missing)
//   network   27 -> 64 -> 64 -> 64 -> 1, ReLU, BF16 weights and activations,
//             FP32 accumulation on MFMA (v_mfma_f32_16x16x32_bf16)
// Training minimises, per sample with residual r to the converged surface,
//   l = c r^2 / (2 s^2) - log(c) / 2        (s = FF_MLP_RES_SCALE)
// whose minimiser is c = s^2 / E[r^2 | features]: inverse-variance weighting.
// At inference a confidence under FF_MLP_GATE (an expected residual above
// ~56 mm, beyond the widest truncation band) rejects the pixel: a lone
// outlier, such as a flying pixel in an occluded region, would otherwise
// leave a surface that no later observation competes with.
//
// MFMA data layout.  A wave holds 16 pixels (samples) along N; hidden units
// run along M.  The result of a 16x16x32 MFMA sits with lane l holding
// sample l % 16 and units 16 t + 4 (l / 16) + r in its 4 registers r, which is
// exactly what the next layer's B operand needs once the K order is permuted:
// K position p = 8 g + q of K-step s stands for unit
//   u(s, p) = 16 (2 s + q / 4) + 4 g + q % 4.
// The weight fragments (A operands) are packed in that order, so the layers
// chain in registers with no LDS transposes.
#include "ff_device.hpp"
#include "ff_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace ff {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define MLP_IN 27
#define MLP_H 64
// parameter layout (fp32 master copy)
#define P_W1 0
#define P_B1 (64 * 27)
#define P_W2 (P_B1 + 64)
#define P_B2 (P_W2 + 4096)
#define P_W3 (P_B2 + 64)
#define P_B3 (P_W3 + 4096)
#define P_W4 (P_B3 + 64)
#define P_B4 (P_W4 + 64)
#define P_COUNT (P_B4 + 1)
static_assert(P_COUNT == FF_MLP_NPARAMS, "parameter layout");

// packed weight fragments: [fragment][lane] x 8 bf16
#define F_A1 0    // W1, 4 M-tiles x 1 K-step
#define F_A2 4    // W2, 4 x 2
#define F_A3 12   // W3, 4 x 2
#define F_A3T 20  // W3 transposed (backward), 4 x 2
#define F_A2T 28  // W2 transposed (backward), 4 x 2
#define F_A4 36   // w4 as row 0 of a 16 x 64 matrix (inference), 1 x 2
#define F_COUNT 38

__host__ __device__ __forceinline__ int unit_of(int s, int p)
{
    const int g = p >> 3, q = p & 7;
    return 16 * (2 * s + (q >> 2)) + 4 * g + (q & 3);
}

__device__ __forceinline__ float relu(float x) { return x > 0.0f ? x : 0.0f; }
// the same for finite x, as an integer max of the bits (a float max of an
// MFMA result would first be canonicalised: one more instruction)
__device__ __forceinline__ float relu_i(float x) { return __int_as_float(max(__float_as_int(x), 0)); }

// ---------------------------------------------------------------------------
// weight packing: fp32 master -> bf16 fragments, biases, w4 (bf16-rounded)
// ---------------------------------------------------------------------------
__global__ void k_mlp_pack(const float *__restrict__ P, bf16x8 *__restrict__ frag, float *__restrict__ bias,
                           float *__restrict__ w4)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F_COUNT * 64) {
        const int f = i / 64, l = i % 64, g = l / 16;
        const int row16 = l % 16;
        bf16x8 v;
        for (int q = 0; q < 8; q++) {
            const int p = 8 * g + q;
            float w = 0.0f;
            if (f < F_A2) {
                const int row = 16 * f + row16, k = p;
                w = k < MLP_IN ? P[P_W1 + row * MLP_IN + k] : 0.0f;
            } else if (f >= F_A4) {
                w = row16 == 0 ? P[P_W4 + unit_of(f - F_A4, p)] : 0.0f;
            } else {
                const int j = (f - F_A2) % 8, t = j / 2, s = j % 2, row = 16 * t + row16, u = unit_of(s, p);
                if (f < F_A3) w = P[P_W2 + row * 64 + u];
                else if (f < F_A3T) w = P[P_W3 + row * 64 + u];
                else if (f < F_A2T) w = P[P_W3 + u * 64 + row];
                else w = P[P_W2 + u * 64 + row];
            }
            v[q] = (__bf16)w;
        }
        frag[i] = v;
    }
    if (i < 3 * 64 + 1) {
        const int src = i < 64 ? P_B1 + i : (i < 128 ? P_B2 + i - 64 : (i < 192 ? P_B3 + i - 128 : P_B4));
        bias[i] = P[src];
    }
    if (i < 64) w4[i] = (float)(__bf16)P[P_W4 + i];
}

struct MlpW {
    bf16x8 a1[4], a2[4][2], a3[4][2];
};

__device__ __forceinline__ void load_forward(MlpW &W, const bf16x8 *__restrict__ frag, int lane)
{
#pragma unroll
    for (int t = 0; t < 4; t++) {
        W.a1[t] = frag[(F_A1 + t) * 64 + lane];
#pragma unroll
        for (int s = 0; s < 2; s++) {
            W.a2[t][s] = frag[(F_A2 + 2 * t + s) * 64 + lane];
            W.a3[t][s] = frag[(F_A3 + 2 * t + s) * 64 + lane];
        }
    }
}

// B operand of the next layer from this layer's 4 result tiles (bf16 values)
__device__ __forceinline__ void to_b(const f32x4 h[4], bf16x8 b[2])
{
#pragma unroll
    for (int s = 0; s < 2; s++) {
#pragma unroll
        for (int q = 0; q < 8; q++) b[s][q] = (__bf16)h[2 * s + (q >> 2)][q & 3];
    }
}

// one dense layer: z = A x B + bias, for 4 M-tiles and nk K-steps
template <int NK>
__device__ __forceinline__ void layer(const bf16x8 (&a)[4][2], const bf16x8 *b, const float *__restrict__ bias,
                                      int g, f32x4 z[4])
{
#pragma unroll
    for (int t = 0; t < 4; t++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < NK; s++) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t][s], b[s], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; r++) z[t][r] = acc[r] + bias[16 * t + 4 * g + r];
    }
}

struct Fwd {
    f32x4 z1[4], z2[4], z3[4];  // pre-activations
    f32x4 h1[4], h2[4], h3[4];  // bf16-rounded activations (as float)
    float o, c;
};

// forward for the wave's 16 samples; bx = features as the first B operand
__device__ __forceinline__ void mlp_forward(const MlpW &W, const bf16x8 &bx, const float *__restrict__ bias,
                                            const float *__restrict__ w4, int lane, Fwd &F)
{
    const int g = lane >> 4;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a1[t], bx, acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            F.z1[t][r] = acc[r] + bias[16 * t + 4 * g + r];
            F.h1[t][r] = (float)(__bf16)relu(F.z1[t][r]);
        }
    }
    bf16x8 b[2];
    to_b(F.h1, b);
    layer<2>(W.a2, b, bias + 64, g, F.z2);
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) F.h2[t][r] = (float)(__bf16)relu(F.z2[t][r]);
    to_b(F.h2, b);
    layer<2>(W.a3, b, bias + 128, g, F.z3);
    float part = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            F.h3[t][r] = (float)(__bf16)relu(F.z3[t][r]);
            part += w4[16 * t + 4 * g + r] * F.h3[t][r];
        }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    F.o = part + bias[192];
    F.c = 1.0f / (1.0f + expf(-F.o));
}

// ---------------------------------------------------------------------------
// features of pixel (x, y); lane group g fills positions 8 g .. 8 g + 7
// ---------------------------------------------------------------------------
struct FeatCam {
    float fx, fy, cx, cy, scaling, maxcam;
    float ifx, ify;  // 1 / fx, 1 / fy (float division on the host)
};

__device__ __forceinline__ float depth_at(const uint16_t *__restrict__ depth, int W, int H, int x, int y,
                                          float scaling)
{
    if (x < 0 || y < 0 || x >= W || y >= H) return 0.0f;
    return __fmul_rn((float)depth[(size_t)y * W + x], scaling);
}

// patch entry: z_k / z_c - 1 with rc = 1 / z_c, clamped to [-1, 1]
__device__ __forceinline__ float patch_feature(float zk, float rc)
{
    return fminf(fmaxf(__fsub_rn(__fmul_rn(zk, rc), 1.0f), -1.0f), 1.0f);
}

// |cos| between the viewing ray of (x, y) and the normal from the central
// differences of the four neighbours' back-projections (0 if one is missing)
__device__ float incidence_cos(float zl, float zr, float zu, float zd, float zc, int x, int y, const FeatCam &K)
{
    if (!(zl > 0.0f && zr > 0.0f && zu > 0.0f && zd > 0.0f)) return 0.0f;
    const float ul = __fmul_rn((float)(x - 1) - K.cx, K.ifx), ur = __fmul_rn((float)(x + 1) - K.cx, K.ifx);
    const float uc = __fmul_rn((float)x - K.cx, K.ifx);
    const float vu = __fmul_rn((float)(y - 1) - K.cy, K.ify), vd = __fmul_rn((float)(y + 1) - K.cy, K.ify);
    const float vc = __fmul_rn((float)y - K.cy, K.ify);
    // tx = P(x+1) - P(x-1), ty = P(y+1) - P(y-1)
    const float tx0 = __fsub_rn(__fmul_rn(ur, zr), __fmul_rn(ul, zl));
    const float tx1 = __fsub_rn(__fmul_rn(vc, zr), __fmul_rn(vc, zl));
    const float tx2 = __fsub_rn(zr, zl);
    const float ty0 = __fsub_rn(__fmul_rn(uc, zd), __fmul_rn(uc, zu));
    const float ty1 = __fsub_rn(__fmul_rn(vd, zd), __fmul_rn(vu, zu));
    const float ty2 = __fsub_rn(zd, zu);
    const float n0 = __fsub_rn(__fmul_rn(tx1, ty2), __fmul_rn(tx2, ty1));
    const float n1 = __fsub_rn(__fmul_rn(tx2, ty0), __fmul_rn(tx0, ty2));
    const float n2 = __fsub_rn(__fmul_rn(tx0, ty1), __fmul_rn(tx1, ty0));
    const float r0 = __fmul_rn(uc, zc), r1 = __fmul_rn(vc, zc), r2 = zc;
    const float nn = __fadd_rn(__fadd_rn(__fmul_rn(n0, n0), __fmul_rn(n1, n1)), __fmul_rn(n2, n2));
    const float rr = __fadd_rn(__fadd_rn(__fmul_rn(r0, r0), __fmul_rn(r1, r1)), __fmul_rn(r2, r2));
    if (!(nn > 0.0f) || !(rr > 0.0f)) return 0.0f;
    const float dot = __fadd_rn(__fadd_rn(__fmul_rn(n0, r0), __fmul_rn(n1, r1)), __fmul_rn(n2, r2));
    return __fdiv_rn(fabsf(dot), __fsqrt_rn(__fmul_rn(nn, rr)));
}

// feature k of pixel (x, y) read from the image (training-data path)
__device__ float feature(const uint16_t *__restrict__ depth, int W, int H, int x, int y, float zc, int k,
                         const FeatCam &K)
{
    if (k < 25) return patch_feature(depth_at(depth, W, H, x + k % 5 - 2, y + k / 5 - 2, K.scaling), __frcp_rn(zc));
    if (k == 25) return __fmul_rn(zc, 0.25f);
    if (k == 26)
        return incidence_cos(depth_at(depth, W, H, x - 1, y, K.scaling), depth_at(depth, W, H, x + 1, y, K.scaling),
                             depth_at(depth, W, H, x, y - 1, K.scaling), depth_at(depth, W, H, x, y + 1, K.scaling), zc,
                             x, y, K);
    return 0.0f;
}

// ---------------------------------------------------------------------------
// inference: per-pixel confidence of a frame (0 where the depth is invalid)
// A wave takes 16 pixels of a row; the 5 x 20 depth window they need is
// staged in the wave's LDS slice, every lane of a pixel (one per 8-feature
// group) computes the same incidence, and feature selection is branch-free.
// ---------------------------------------------------------------------------
// A wave takes 64 pixels of a row per step: the 5 x 68 depth window they
// need is staged in the wave's LDS slice; each lane computes its own pixel's
// centre depth, validity, reciprocal and incidence once, and then, for each
// of the four 16-pixel MFMA tiles, its 8 features of one pixel of the tile
// (feature selection branch-free).  The network runs once per tile.
#define CONF_WIN_W 68
#define CONF_WIN (5 * CONF_WIN_W)

// inference forward of mlp_forward: each layer's accumulator starts at its
// bias (the MFMA adds the products to it), the activations go straight to
// the next layer's bf16 operand, and the output layer (w4, bf16 as in
// mlp_forward) is two more MFMAs with w4 as row 0 of the A matrix, so its
// result is row 0 of the tile: lanes 0-15 (g = 0) hold their pixel's o.
// Only the summation order differs from mlp_forward.
__device__ __forceinline__ float mlp_infer(const MlpW &W, const bf16x8 (&a4)[2], const bf16x8 &bx,
                                           const float *__restrict__ bias, int g)
{
    bf16x8 b[2];
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
#pragma unroll
        for (int r = 0; r < 4; r++) acc[t][r] = bias[16 * t + 4 * g + r];
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a1[t], bx, acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int q = 0; q < 8; q++) b[s][q] = (__bf16)relu_i(acc[2 * s + (q >> 2)][q & 3]);
#pragma unroll
    for (int t = 0; t < 4; t++) {
#pragma unroll
        for (int r = 0; r < 4; r++) acc[t][r] = bias[64 + 16 * t + 4 * g + r];
#pragma unroll
        for (int s = 0; s < 2; s++)
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a2[t][s], b[s], acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int q = 0; q < 8; q++) b[s][q] = (__bf16)relu_i(acc[2 * s + (q >> 2)][q & 3]);
#pragma unroll
    for (int t = 0; t < 4; t++) {
#pragma unroll
        for (int r = 0; r < 4; r++) acc[t][r] = bias[128 + 16 * t + 4 * g + r];
#pragma unroll
        for (int s = 0; s < 2; s++)
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a3[t][s], b[s], acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int q = 0; q < 8; q++) b[s][q] = (__bf16)relu_i(acc[2 * s + (q >> 2)][q & 3]);
    f32x4 out = {bias[192], 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 2; s++) out = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a4[s], b[s], out, 0, 0, 0);
    return 1.0f / (1.0f + expf(-out[0]));
}

struct ConfWave {
    float z[CONF_WIN];
    float zc[64], rc[64], inc[64];
};

__device__ __forceinline__ void confidence_body(const uint16_t *__restrict__ depth, const FeatCam K, int W, int H,
                                                const bf16x8 *__restrict__ frag, const float *__restrict__ bias,
                                                float *__restrict__ conf)
{
    __shared__ float sb[3 * 64 + 1];
    __shared__ ConfWave cw[4];
    for (int i = threadIdx.x; i < 3 * 64 + 1; i += blockDim.x) sb[i] = bias[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4, p = lane & 15;
    ConfWave &C = cw[threadIdx.x >> 6];
    MlpW Wt;
    load_forward(Wt, frag, lane);
    const bf16x8 a4[2] = {frag[F_A4 * 64 + lane], frag[(F_A4 + 1) * 64 + lane]};
    const int segsPerRow = (W + 63) / 64;
    const int nSegs = segsPerRow * H;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nWaves = (gridDim.x * blockDim.x) >> 6;
    for (int seg = wave; seg < nSegs; seg += nWaves) {
        const int y = seg / segsPerRow, x0 = (seg % segsPerRow) * 64;
        // window rows y-2..y+2, columns x0-2..x0+65
#pragma unroll
        for (int k = 0; k < (CONF_WIN + 63) / 64; k++) {
            const int i = lane + 64 * k;
            if (i < CONF_WIN) C.z[i] = depth_at(depth, W, H, x0 - 2 + i % CONF_WIN_W, y - 2 + i / CONF_WIN_W, K.scaling);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        {
            // this lane's own pixel x0 + lane
            const int x = x0 + lane;
            const float zc = C.z[2 * CONF_WIN_W + lane + 2];
            const bool valid = x < W && zc > 0.0f && zc < K.maxcam;
            C.zc[lane] = valid ? zc : -1.0f;  // -1: not a valid pixel
            C.rc[lane] = __frcp_rn(valid ? zc : 1.0f);
            C.inc[lane] = incidence_cos(C.z[2 * CONF_WIN_W + lane + 1], C.z[2 * CONF_WIN_W + lane + 3],
                                        C.z[CONF_WIN_W + lane + 2], C.z[3 * CONF_WIN_W + lane + 2], zc, x, y, K);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 1
        for (int j = 0; j < 4; j++) {
            const int px = 16 * j + p;  // the tile's pixel of this lane
            const float zc = C.zc[px], rc = C.rc[px], inc = C.inc[px];
            const bool valid = zc >= 0.0f;
            bf16x8 bx;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int k = 8 * g + q, kk = k < 25 ? k : 0;
                const float fp = patch_feature(C.z[(kk / 5) * CONF_WIN_W + px + kk % 5], rc);
                const float f = k < 25 ? fp : (k == 25 ? __fmul_rn(zc, 0.25f) : (k == 26 ? inc : 0.0f));
                bx[q] = (__bf16)(valid ? f : 0.0f);
            }
            // the biases are re-read from LDS for every tile (held in
            // registers across tiles they would cost the occupancy)
            const float *bsb = sb;
            asm volatile("" : "+v"(bsb));
            const float cc = mlp_infer(Wt, a4, bx, bsb, g);
            // a pixel whose predicted residual is beyond every truncation band
            // (confidence under the gate) is not fused at all
            const int x = x0 + px;
            if (g == 0 && x < W) conf[(size_t)y * W + x] = (valid && cc >= FF_MLP_GATE) ? cc : 0.0f;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// the frame's values from the device frame block (a captured frame replays it)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_confidence(const FrameDev *__restrict__ fd, int W, int H,
                                                    const bf16x8 *__restrict__ frag, const float *__restrict__ bias,
                                                    const float *__restrict__ w4, float *__restrict__ conf)
{
    const FeatCam K{fd->sp.fx, fd->sp.fy, fd->sp.cx, fd->sp.cy, fd->scaling, fd->maxcam, fd->ifx, fd->ify};
    (void)w4;
    confidence_body(fd->depth, K, W, H, frag, bias, conf);
}

// the frame's values as arguments (direct launches beside the frame's
// allocation, before its frame block is on the device)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_confidence_v(const uint16_t *__restrict__ depth, FeatCam K, int W, int H,
                                                      const bf16x8 *__restrict__ frag, const float *__restrict__ bias,
                                                      const float *__restrict__ w4, float *__restrict__ conf)
{
    (void)w4;
    confidence_body(depth, K, W, H, frag, bias, conf);
}

// ---------------------------------------------------------------------------
// training data: features + residual to a converged map, per valid pixel
// ---------------------------------------------------------------------------
struct MapRead {
    const uint32_t *tree;
    const ushort4 *info;
    const uint32_t *parent;
    const float *dist, *wgt;
    uint32_t nLeavesTotal, n;
    float off[3], scale, trunc;
};

// trilinear TSDF at voxel coordinates v of the finest brick with weight
// there, and that brick's truncation (distance threshold x leaf scale)
__device__ bool tsdf_at(const MapRead &M, const float v[3], float &d, float &td)
{
    if (!(v[0] >= 0.0f && v[1] >= 0.0f && v[2] >= 0.0f && v[0] < (float)M.n && v[1] < (float)M.n &&
          v[2] < (float)M.n))
        return false;
    const int X = (int)v[0], Y = (int)v[1], Z = (int)v[2];
    uint32_t half = M.n >> 1, branch = 0, leaf = FF_BRANCHINIT;
    int ox = 0, oy = 0, oz = 0;
    for (;;) {
        uint32_t co = 0;
        if (X >= ox + (int)half) { co |= 1u; ox += half; }
        if (Y >= oy + (int)half) { co |= 2u; oy += half; }
        if (Z >= oz + (int)half) { co |= 4u; oz += half; }
        const uint32_t l = M.tree[branch + co];
        if (l < M.nLeavesTotal) leaf = l;
        const uint32_t child = M.tree[branch + co + 8];
        if (child == FF_BRANCHINIT || half <= FF_BL) break;
        branch = child;
        half >>= 1;
    }
    for (int guard = 0; leaf < FF_BRANCHINIT && guard < 16; leaf = M.parent[leaf], guard++) {
        const ushort4 li = M.info[leaf];
        const float ls = (float)li.w;
        float u[3];
        int i0[3], i1[3];
        float fr[3];
        const float o[3] = {(float)li.x, (float)li.y, (float)li.z};
        for (int k = 0; k < 3; k++) {
            u[k] = fminf(fmaxf((v[k] - o[k]) / ls, 0.0f), 7.0f);
            i0[k] = min((int)u[k], 6);
            i1[k] = i0[k] + 1;
            fr[k] = u[k] - (float)i0[k];
        }
        float acc = 0.0f, wsum = 0.0f;
        bool ok = true;
        for (int c = 0; c < 8; c++) {
            const int ix = (c & 1) ? i1[0] : i0[0], iy = (c & 2) ? i1[1] : i0[1], iz = (c & 4) ? i1[2] : i0[2];
            const size_t q = (size_t)leaf * FF_BRICK + (iz * FF_BL + iy) * FF_BL + ix;
            const float w = M.wgt[q];
            if (!(w > 0.0f)) ok = false;
            const float f = ((c & 1) ? fr[0] : 1.0f - fr[0]) * ((c & 2) ? fr[1] : 1.0f - fr[1]) *
                            ((c & 4) ? fr[2] : 1.0f - fr[2]);
            acc += f * M.dist[q];
            wsum += f;
        }
        if (ok) {
            d = acc / wsum;
            td = M.trunc * ls;
            return true;
        }
    }
    return false;
}

__device__ __forceinline__ uint32_t hash32(uint32_t a, uint32_t b)
{
    uint32_t h = a * 0x9e3779b1u ^ (b + 0x7f4a7c15u + (a << 6) + (a >> 2));
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h;
}

struct CollectOut {
    uint16_t *feat;   // [cap][32] bf16 bits
    float *resid;     // [cap]
    uint32_t *count;
    uint32_t cap;
};

__global__ __launch_bounds__(256) void k_collect(const uint16_t *__restrict__ depth, int W, int H, FeatCam K,
                                                 CamF pw, MapRead M, uint32_t keepThresh, uint32_t seed,
                                                 CollectOut out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    const int lane = threadIdx.x & 63;
    bool take = false;
    float zc = 0.0f, r = 0.0f;
    if (x < W) {
        zc = depth_at(depth, W, H, x, y, K.scaling);
        take = zc > 0.0f && zc < K.maxcam && hash32((uint32_t)(y * W + x), seed) < keepThresh;
        if (take) {
            // camera point -> world -> voxel coordinates of the map
            const float cxm = (float)(x - K.cx) / K.fx * zc, cym = (float)(y - K.cy) / K.fy * zc;
            float v[3];
            for (int k = 0; k < 3; k++) {
                const float wv = pw.r[3 * k] * cxm + pw.r[3 * k + 1] * cym + pw.r[3 * k + 2] * zc + pw.t[k];
                v[k] = (wv - M.off[k]) / M.scale;
            }
            // residual to the converged surface: the TSDF inside its band; a
            // point the band does not reach (saturated, or never observed)
            // is an outlier at FF_MLP_RES_OUT
            float d, td;
            if (tsdf_at(M, v, d, td))
                r = fabsf(d) < 0.9f * td ? d : (d < 0.0f ? -FF_MLP_RES_OUT : FF_MLP_RES_OUT);
            else
                r = FF_MLP_RES_OUT;
        }
    }
    const uint64_t m = __ballot(take);
    if (!m) return;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(out.count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (!take) return;
    const uint32_t i = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (i >= out.cap) return;
    for (int k = 0; k < 32; k++) {
        const __bf16 b = (__bf16)feature(depth, W, H, x, y, zc, k, K);
        out.feat[(size_t)i * 32 + k] = __builtin_bit_cast(uint16_t, b);
    }
    out.resid[i] = r;
}

// ---------------------------------------------------------------------------
// training: T1 forward + backward per 16-sample tile, T2 weight gradients
// (split over samples), T3 reduction + Adam + repacking
// ---------------------------------------------------------------------------
struct Batch {
    const bf16x8 *feat;     // dataset features, [n][4] bf16x8
    const float *resid;     // dataset residuals
    const uint32_t *index;  // optional explicit sample indices (gradient checks)
    uint32_t nData, B, step, seed;
    float invB, invS2;
};

struct Acts {  // [rows][B] bf16 matrices and the output gradient
    __bf16 *X, *H1, *H2, *H3, *D1, *D2, *D3;
    float *G;
    float *loss;  // per-block partial losses
};

__device__ __forceinline__ void store_tiles(__bf16 *M, const f32x4 v[4], int g, int col, uint32_t B)
{
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) M[(size_t)(16 * t + 4 * g + r) * B + col] = (__bf16)v[t][r];
}

__global__ __launch_bounds__(256) void k_train_fb(Batch bt, const bf16x8 *__restrict__ frag,
                                                  const float *__restrict__ bias, const float *__restrict__ w4,
                                                  Acts A)
{
    __shared__ bf16x8 sf[F_A4 * 64];  // the fragments the training pass uses
    __shared__ float sb[3 * 64 + 1], sw4[64], sl[4];
    for (int i = threadIdx.x; i < F_A4 * 64; i += blockDim.x) sf[i] = frag[i];
    for (int i = threadIdx.x; i < 3 * 64 + 1; i += blockDim.x) sb[i] = bias[i];
    for (int i = threadIdx.x; i < 64; i += blockDim.x) sw4[i] = w4[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4, wv = threadIdx.x >> 6;
    MlpW Wt;
    load_forward(Wt, sf, lane);
    const uint32_t nTiles = bt.B / 16;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nWaves = (gridDim.x * blockDim.x) >> 6;
    float lossAcc = 0.0f;
    for (uint32_t tile = wave; tile < nTiles; tile += nWaves) {
        const uint32_t col = tile * 16 + (lane & 15);
        const uint32_t idx = bt.index ? bt.index[col] : hash32(bt.step * 0x10001u + col, bt.seed) % bt.nData;
        const bf16x8 bx = bt.feat[(size_t)idx * 4 + g];
        const float r = bt.resid[idx];
        Fwd F;
        mlp_forward(Wt, bx, sb, sw4, lane, F);
        // l = c r^2 / (2 s^2) - log(c) / 2 ;  dl/do = (1 - c) (c r^2 / (2 s^2) - 1/2)
        const float r2s = 0.5f * r * r * bt.invS2;
        const float softplus = F.o > 0.0f ? F.o + log1pf(expf(-F.o)) : log1pf(expf(F.o));  // -log(1 - c) ... log(1+e^o)
        const float logc = F.o - softplus;
        if (g == 0) lossAcc += (F.c * r2s - 0.5f * logc) * bt.invB;
        const float go = (1.0f - F.c) * (F.c * r2s - 0.5f) * bt.invB;
        // delta3 = go * w4 * relu'(z3)
        f32x4 d3[4], d2[4], d1[4];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                d3[t][r] = (float)(__bf16)(F.z3[t][r] > 0.0f ? go * sw4[16 * t + 4 * g + r] : 0.0f);
        bf16x8 b[2];
        to_b(d3, b);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < 2; s++)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sf[(F_A3T + 2 * t + s) * 64 + lane], b[s], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) d2[t][r] = (float)(__bf16)(F.z2[t][r] > 0.0f ? acc[r] : 0.0f);
        }
        to_b(d2, b);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < 2; s++)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sf[(F_A2T + 2 * t + s) * 64 + lane], b[s], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) d1[t][r] = (float)(__bf16)(F.z1[t][r] > 0.0f ? acc[r] : 0.0f);
        }
        // the matrices the weight gradients contract over samples
#pragma unroll
        for (int q = 0; q < 8; q++) A.X[(size_t)(8 * g + q) * bt.B + col] = bx[q];
        store_tiles(A.H1, F.h1, g, col, bt.B);
        store_tiles(A.H2, F.h2, g, col, bt.B);
        store_tiles(A.H3, F.h3, g, col, bt.B);
        store_tiles(A.D1, d1, g, col, bt.B);
        store_tiles(A.D2, d2, g, col, bt.B);
        store_tiles(A.D3, d3, g, col, bt.B);
        if (g == 0) A.G[col] = go;
    }
    for (int o = 32; o > 0; o >>= 1) lossAcc += __shfl_xor(lossAcc, o);
    if (lane == 0) sl[wv] = lossAcc;
    __syncthreads();
    if (threadIdx.x == 0) A.loss[blockIdx.x] = (sl[0] + sl[1]) + (sl[2] + sl[3]);
}

// T2: per block, partial gradients over samples [k0, k0 + KB) into part[block][P_COUNT]
#define T2_KB 2048
__global__ __launch_bounds__(256) void k_train_wgrad(Acts A, uint32_t B, float *__restrict__ part)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, wv = threadIdx.x >> 6;
    const uint32_t k0 = blockIdx.x * T2_KB;
    float *out = part + (size_t)blockIdx.x * P_COUNT;
    // 40 output tiles: dW1 4 x 2 (ids 0-7), dW2 4 x 4 (8-23), dW3 4 x 4 (24-39)
    for (int id = wv; id < 40; id += 4) {
        const __bf16 *Dm, *Hm;
        int tm, tn, base, ncol, nvalid;
        if (id < 8) { Dm = A.D1; Hm = A.X; tm = id / 2; tn = id % 2; base = P_W1; ncol = MLP_IN; nvalid = MLP_IN; }
        else if (id < 24) { Dm = A.D2; Hm = A.H1; tm = (id - 8) / 4; tn = (id - 8) % 4; base = P_W2; ncol = 64; nvalid = 64; }
        else { Dm = A.D3; Hm = A.H2; tm = (id - 24) / 4; tn = (id - 24) % 4; base = P_W3; ncol = 64; nvalid = 64; }
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const __bf16 *arow = Dm + (size_t)(16 * tm + (lane & 15)) * B + k0 + 8 * g;
        const __bf16 *brow = Hm + (size_t)(16 * tn + (lane & 15)) * B + k0 + 8 * g;
        for (uint32_t k = 0; k < T2_KB; k += 32) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(arow + k);
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(brow + k);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        // D[i][j]: lane -> j = 16 tn + lane % 16, reg r -> i = 16 tm + 4 g + r
        const int j = 16 * tn + (lane & 15);
        if (j < nvalid)
#pragma unroll
            for (int r = 0; r < 4; r++) out[base + (16 * tm + 4 * g + r) * ncol + j] = acc[r];
    }
    // bias gradients, w4 and b4 gradients: row sums over the chunk
    const int t = threadIdx.x;
    if (t < 192) {
        const __bf16 *row = (t < 64 ? A.D1 : (t < 128 ? A.D2 : A.D3)) + (size_t)(t & 63) * B + k0;
        float s = 0.0f;
        for (uint32_t k = 0; k < T2_KB; k++) s += (float)row[k];
        out[(t < 64 ? P_B1 : (t < 128 ? P_B2 : P_B3)) + (t & 63)] = s;
    } else {
        const int u = t - 192;
        const __bf16 *row = A.H3 + (size_t)u * B + k0;
        float s = 0.0f, sg = 0.0f;
        for (uint32_t k = 0; k < T2_KB; k++) {
            const float gk = A.G[k0 + k];
            s += gk * (float)row[k];
            sg += gk;
        }
        out[P_W4 + u] = s;
        if (u == 0) out[P_B4] = sg;
    }
}

struct Adam {
    float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
    int update;                     // 0: only reduce the gradient
};

__global__ void k_train_update(const float *__restrict__ part, int nPart, float *__restrict__ grad,
                               float *__restrict__ P, float *__restrict__ m1, float *__restrict__ m2, Adam ad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P_COUNT) return;
    float gsum = 0.0f;
    for (int b = 0; b < nPart; b++) gsum += part[(size_t)b * P_COUNT + i];
    grad[i] = gsum;
    if (!ad.update) return;
    const float m = ad.b1 * m1[i] + (1.0f - ad.b1) * gsum;
    const float v = ad.b2 * m2[i] + (1.0f - ad.b2) * gsum * gsum;
    m1[i] = m;
    m2[i] = v;
    P[i] -= ad.lr * (m / ad.c1) / (sqrtf(v / ad.c2) + ad.eps);
}

// ---------------------------------------------------------------------------
// state
// ---------------------------------------------------------------------------
struct MlpBuffers {
    float *P = nullptr;       // fp32 master
    bf16x8 *frag = nullptr;
    float *bias = nullptr, *w4 = nullptr;
    int alloc(hipStream_t st)
    {
        if (hipMalloc(&P, sizeof(float) * P_COUNT) != hipSuccess) return -1;
        if (hipMalloc(&frag, sizeof(bf16x8) * F_COUNT * 64) != hipSuccess) return -1;
        if (hipMalloc(&bias, sizeof(float) * (3 * 64 + 1)) != hipSuccess) return -1;
        if (hipMalloc(&w4, sizeof(float) * 64) != hipSuccess) return -1;
        (void)st;
        return 0;
    }
    void free_all()
    {
        hipFree(P); hipFree(frag); hipFree(bias); hipFree(w4);
        P = nullptr; frag = nullptr; bias = w4 = nullptr;
    }
    void pack(hipStream_t st) { k_mlp_pack<<<(F_COUNT * 64 + 255) / 256, 256, 0, st>>>(P, frag, bias, w4); }
};

struct LearnedState {
    MlpBuffers w;
    // the buffer the current frame's confidence pass writes: one of bufs,
    // or a caller's buffer swapped in for one frame (external)
    float *conf = nullptr;
    size_t confCap = 0;
    // overlapped frames alternate between two buffers: a frame's pass then
    // writes the buffer the frame before last read, whose SDF update has
    // ended (it precedes the previous frame's allocation on the side
    // stream, whose report the host awaited), so the pass need not wait for
    // the previous frame's SDF update, which reads the other buffer
    float *bufs[2] = {nullptr, nullptr};
    size_t bufCap[2] = {0, 0};
    int cur = 0;
    bool external = false;
};

static bool lok(ff_ctx *c, hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

void learned_free(ff_ctx *c)
{
    if (!c->learned) return;
    c->bufGen++;
    c->learned->w.free_all();
    hipFree(c->learned->bufs[0]);
    hipFree(c->learned->bufs[1]);
    delete c->learned;
    c->learned = nullptr;
    c->lastConf = nullptr;
}

// the confidence buffer for a W x H frame, or nullptr on failure;
// alternate: take the other of the two buffers (an overlapped frame)
const float *learned_prepare(ff_ctx *c, int W, int H, bool alternate)
{
    LearnedState *L = c->learned;
    if (!L) return nullptr;
    const size_t np = (size_t)W * H;
    if (L->external) return L->confCap >= np ? L->conf : nullptr;
    if (alternate) {
        L->cur ^= 1;
        c->bufGen++;  // (a captured graph holds the other buffer)
    }
    const int k = L->cur;
    if (L->bufCap[k] < np) {
        frame_join(c);  // the frame in flight may still read the confidences
        hipFree(L->bufs[k]);
        L->bufs[k] = nullptr;
        L->bufCap[k] = 0;
        if (!lok(c, hipMalloc(&L->bufs[k], sizeof(float) * np), "hipMalloc(confidence)")) return nullptr;
        L->bufCap[k] = np;
        c->bufGen++;
    }
    L->conf = L->bufs[k];
    L->confCap = L->bufCap[k];
    return L->conf;
}

// the confidence buffer replaced by (buf, cap) for one frame; the previous
// pair goes to (*old, *oldCap) when those are given (no-op with learned off)
void learned_swap_conf(ff_ctx *c, float *buf, size_t cap, float **old, size_t *oldCap)
{
    LearnedState *L = c->learned;
    if (!L) return;
    if (old) *old = L->conf;
    if (oldCap) *oldCap = L->confCap;
    L->conf = buf;
    L->confCap = cap;
    L->external = buf != L->bufs[0] && buf != L->bufs[1];
}

// per-pixel confidence of the frame described by c->d_fd
void learned_launch(ff_ctx *c, int W, int H, hipStream_t st)
{
    LearnedState *L = c->learned;
    const int segs = (W + 63) / 64 * H;
    const int blocks = std::min(2048, (segs + 3) / 4);
    k_confidence<<<blocks, 256, 0, st>>>(c->d_fd, W, H, L->w.frag, L->w.bias, L->w.w4, L->conf);
}

// the same with the frame's values from the host frame block (c->h_fd)
void learned_launch_direct(ff_ctx *c, int W, int H, hipStream_t st)
{
    LearnedState *L = c->learned;
    const FrameDev &f = *c->h_fd;
    const FeatCam K{f.sp.fx, f.sp.fy, f.sp.cx, f.sp.cy, f.scaling, f.maxcam, f.ifx, f.ify};
    const int segs = (W + 63) / 64 * H;
    const int blocks = std::min(2048, (segs + 3) / 4);
    k_confidence_v<<<blocks, 256, 0, st>>>(f.depth, K, W, H, L->w.frag, L->w.bias, L->w.w4, L->conf);
}

}  // namespace ff

using namespace ff;

// ---------------------------------------------------------------------------
// trainer
// ---------------------------------------------------------------------------
struct ff_trainer {
    int device = 0;
    hipStream_t stream = nullptr;
    MlpBuffers w;
    float *m1 = nullptr, *m2 = nullptr, *grad = nullptr, *part = nullptr, *lossPart = nullptr;
    uint16_t *feat = nullptr;
    float *resid = nullptr;
    uint32_t *count = nullptr, *index = nullptr;
    size_t cap = 0, n = 0;
    uint32_t batchCap = 0;
    Acts acts{};
    uint32_t step = 0;
    uint64_t seed = 0;
    std::string err;
};

static bool tok(ff_trainer *t, hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    t->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}
#define TCHK(call)                                      \
    do {                                                \
        if (!tok(tr, (call), #call)) return FF_E_DEVICE; \
    } while (0)

static int ensure_batch(ff_trainer *tr, uint32_t B)
{
    if (tr->batchCap >= B) return FF_OK;
    Acts &a = tr->acts;
    hipFree(a.X); hipFree(a.H1); hipFree(a.H2); hipFree(a.H3); hipFree(a.D1); hipFree(a.D2); hipFree(a.D3);
    hipFree(a.G); hipFree(a.loss); hipFree(tr->part); hipFree(tr->index);
    a = Acts{};
    tr->part = nullptr;
    tr->index = nullptr;
    tr->batchCap = 0;
    TCHK(hipMalloc(&a.X, sizeof(__bf16) * 32 * (size_t)B));
    __bf16 **mats[6] = {&a.H1, &a.H2, &a.H3, &a.D1, &a.D2, &a.D3};
    for (int i = 0; i < 6; i++) TCHK(hipMalloc(mats[i], sizeof(__bf16) * 64 * (size_t)B));
    TCHK(hipMalloc(&a.G, sizeof(float) * B));
    TCHK(hipMalloc(&a.loss, sizeof(float) * 4096));
    TCHK(hipMalloc(&tr->part, sizeof(float) * P_COUNT * (size_t)(B / T2_KB)));
    TCHK(hipMalloc(&tr->index, sizeof(uint32_t) * B));
    tr->batchCap = B;
    return FF_OK;
}

// one batch: forward/backward, gradient into tr->grad, optional Adam update
static int train_batch(ff_trainer *tr, uint32_t B, const uint32_t *index, float lr, bool update, float *loss)
{
    int rc = ensure_batch(tr, B);
    if (rc) return rc;
    Batch bt;
    bt.feat = reinterpret_cast<const bf16x8 *>(tr->feat);
    bt.resid = tr->resid;
    bt.index = index;
    bt.nData = (uint32_t)tr->n;
    bt.B = B;
    bt.step = tr->step;
    bt.seed = (uint32_t)tr->seed;
    bt.invB = 1.0f / (float)B;
    bt.invS2 = 1.0f / (FF_MLP_RES_SCALE * FF_MLP_RES_SCALE);
    const int fbBlocks = (int)std::min<uint32_t>(1024u, B / 64);
    k_train_fb<<<fbBlocks, 256, 0, tr->stream>>>(bt, tr->w.frag, tr->w.bias, tr->w.w4, tr->acts);
    const int nPart = (int)(B / T2_KB);
    k_train_wgrad<<<nPart, 256, 0, tr->stream>>>(tr->acts, B, tr->part);
    Adam ad;
    ad.lr = lr; ad.b1 = 0.9f; ad.b2 = 0.999f; ad.eps = 1e-8f;
    ad.update = update ? 1 : 0;
    const double t = (double)tr->step + 1.0;
    ad.c1 = (float)(1.0 - std::pow(0.9, t));
    ad.c2 = (float)(1.0 - std::pow(0.999, t));
    k_train_update<<<(P_COUNT + 255) / 256, 256, 0, tr->stream>>>(tr->part, nPart, tr->grad, tr->w.P, tr->m1, tr->m2,
                                                                  ad);
    if (update) {
        tr->w.pack(tr->stream);
        tr->step++;
    }
    TCHK(hipGetLastError());
    if (loss) {
        std::vector<float> lp(fbBlocks);
        TCHK(hipMemcpyAsync(lp.data(), tr->acts.loss, sizeof(float) * fbBlocks, hipMemcpyDeviceToHost, tr->stream));
        TCHK(hipStreamSynchronize(tr->stream));
        double s = 0.0;
        for (float v : lp) s += v;
        *loss = (float)s;
    }
    return FF_OK;
}

extern "C" {

int ff_set_learned_weighting(ff_ctx *c, const float *params, size_t n)
{
    if (!c) return FF_E_ARG;
    hipSetDevice(c->device);
    if (!params) {
        learned_free(c);
        return FF_OK;
    }
    if (n != P_COUNT) {
        c->err = "ff_set_learned_weighting: expected FF_MLP_NPARAMS parameters";
        return FF_E_ARG;
    }
    if (!c->learned) {
        c->bufGen++;
        c->learned = new LearnedState();
        if (c->learned->w.alloc(c->stream)) {
            learned_free(c);
            c->err = "ff_set_learned_weighting: out of device memory";
            return FF_E_DEVICE;
        }
    }
    if (!lok(c, hipMemcpyAsync(c->learned->w.P, params, sizeof(float) * P_COUNT, hipMemcpyHostToDevice, c->stream),
             "hipMemcpyAsync(params)"))
        return FF_E_DEVICE;
    c->learned->w.pack(c->stream);
    frame_join(c);
    if (!lok(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return FF_E_DEVICE;
    return FF_OK;
}

int ff_copy_confidence(ff_ctx *c, float *out, size_t n)
{
    if (!c || !out) return FF_E_ARG;
    if (!c->learned || !c->lastConf) {
        c->err = "learned weighting is off";
        return FF_E_STATE;
    }
    const size_t np = (size_t)c->imgW * c->imgH;
    if (n < np) return FF_E_ARG;
    frame_join(c);
    if (!lok(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return FF_E_DEVICE;
    if (!lok(c, hipMemcpy(out, c->lastConf, sizeof(float) * np, hipMemcpyDeviceToHost), "hipMemcpy")) return FF_E_DEVICE;
    return FF_OK;
}

ff_trainer *ff_trainer_create(int device, size_t capacity, const float *init, size_t n_params, uint64_t seed)
{
    if (!init || n_params != P_COUNT || capacity == 0 || capacity > (1u << 30)) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    ff_trainer *tr = new ff_trainer();
    tr->device = device;
    tr->seed = seed;
    tr->cap = capacity;
    bool ok = hipStreamCreate(&tr->stream) == hipSuccess && tr->w.alloc(tr->stream) == 0 &&
              hipMalloc(&tr->m1, sizeof(float) * P_COUNT) == hipSuccess &&
              hipMalloc(&tr->m2, sizeof(float) * P_COUNT) == hipSuccess &&
              hipMalloc(&tr->grad, sizeof(float) * P_COUNT) == hipSuccess &&
              hipMalloc(&tr->feat, sizeof(uint16_t) * 32 * capacity) == hipSuccess &&
              hipMalloc(&tr->resid, sizeof(float) * capacity) == hipSuccess &&
              hipMalloc(&tr->count, sizeof(uint32_t)) == hipSuccess;
    ok = ok && hipMemset(tr->m1, 0, sizeof(float) * P_COUNT) == hipSuccess &&
         hipMemset(tr->m2, 0, sizeof(float) * P_COUNT) == hipSuccess &&
         hipMemset(tr->count, 0, sizeof(uint32_t)) == hipSuccess &&
         hipMemcpy(tr->w.P, init, sizeof(float) * P_COUNT, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        ff_trainer_destroy(tr);
        return nullptr;
    }
    tr->w.pack(tr->stream);
    hipStreamSynchronize(tr->stream);
    return tr;
}

void ff_trainer_destroy(ff_trainer *tr)
{
    if (!tr) return;
    hipSetDevice(tr->device);
    if (tr->stream) hipStreamSynchronize(tr->stream);
    tr->w.free_all();
    hipFree(tr->m1); hipFree(tr->m2); hipFree(tr->grad); hipFree(tr->feat); hipFree(tr->resid); hipFree(tr->count);
    Acts &a = tr->acts;
    hipFree(a.X); hipFree(a.H1); hipFree(a.H2); hipFree(a.H3); hipFree(a.D1); hipFree(a.D2); hipFree(a.D3);
    hipFree(a.G); hipFree(a.loss); hipFree(tr->part); hipFree(tr->index);
    if (tr->stream) hipStreamDestroy(tr->stream);
    delete tr;
}

const char *ff_trainer_last_error(const ff_trainer *tr) { return tr ? tr->err.c_str() : "null trainer"; }

int ff_trainer_collect(ff_trainer *tr, ff_ctx *map, const uint16_t *depth, int W, int H, const double R[9],
                       const double t[3], const double K[4], float scaling, float maxcam, float keep)
{
    if (!tr || !map || !depth || !R || !t || !K || W <= 0 || H <= 0 || !(keep > 0.0f && keep <= 1.0f))
        return FF_E_ARG;
    if (!map->n) {
        tr->err = "the map has no volume yet";
        return FF_E_STATE;
    }
    hipSetDevice(tr->device);
    uint16_t *d_depth = nullptr;
    TCHK(hipMalloc(&d_depth, sizeof(uint16_t) * (size_t)W * H));
    hipError_t e = hipMemcpyAsync(d_depth, depth, sizeof(uint16_t) * (size_t)W * H, hipMemcpyHostToDevice, tr->stream);
    frame_join(map);
    if (e == hipSuccess) e = hipStreamSynchronize(map->stream);  // the map is complete
    uint32_t before = 0, after = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&before, tr->count, 4, hipMemcpyDeviceToHost, tr->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(tr->stream);
    if (e == hipSuccess) {
        FeatCam fk{(float)K[0], (float)K[1], (float)K[2], (float)K[3], scaling, maxcam, 1.0f / (float)K[0],
                   1.0f / (float)K[1]};
        CamF pw;
        for (int i = 0; i < 9; i++) pw.r[i] = (float)R[i];
        for (int i = 0; i < 3; i++) pw.t[i] = (float)t[i];
        pw.fx = fk.fx; pw.fy = fk.fy; pw.cx = fk.cx; pw.cy = fk.cy;
        MapRead M;
        M.tree = map->d_tree; M.info = map->d_leafinfo; M.parent = map->d_parent;
        M.dist = map->d_dist; M.wgt = map->d_wgt;
        M.nLeavesTotal = map->nLeavesTotal; M.n = map->n;
        for (int k = 0; k < 3; k++) M.off[k] = map->offset[k];
        M.scale = map->scale;
        M.trunc = map->dT;
        CollectOut out{tr->feat, tr->resid, tr->count, (uint32_t)tr->cap};
        const uint32_t thresh = keep >= 1.0f ? 0xffffffffu : (uint32_t)((double)keep * 4294967296.0);
        k_collect<<<dim3((W + 255) / 256, H), 256, 0, tr->stream>>>(d_depth, W, H, fk, pw, M, thresh,
                                                                     (uint32_t)(tr->seed * 2654435761u + tr->n), out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&after, tr->count, 4, hipMemcpyDeviceToHost, tr->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(tr->stream);
    hipFree(d_depth);
    TCHK(e);
    after = std::min<uint32_t>(after, (uint32_t)tr->cap);
    if (after >= tr->cap) {
        tr->n = tr->cap;
        uint32_t capv = (uint32_t)tr->cap;
        TCHK(hipMemcpy(tr->count, &capv, 4, hipMemcpyHostToDevice));
    } else {
        tr->n = after;
    }
    return (int)(tr->n - std::min<size_t>(before, tr->n));
}

int ff_trainer_size(const ff_trainer *tr, size_t *n)
{
    if (!tr || !n) return FF_E_ARG;
    *n = tr->n;
    return FF_OK;
}

int ff_trainer_copy_samples(ff_trainer *tr, uint16_t *feat, float *resid, size_t n)
{
    if (!tr || n > tr->n) return FF_E_ARG;
    hipSetDevice(tr->device);
    TCHK(hipStreamSynchronize(tr->stream));
    if (feat) TCHK(hipMemcpy(feat, tr->feat, sizeof(uint16_t) * 32 * n, hipMemcpyDeviceToHost));
    if (resid) TCHK(hipMemcpy(resid, tr->resid, sizeof(float) * n, hipMemcpyDeviceToHost));
    return FF_OK;
}

int ff_trainer_set_samples(ff_trainer *tr, const uint16_t *feat, const float *resid, size_t n)
{
    if (!tr || !feat || !resid || n == 0 || n > tr->cap) return FF_E_ARG;
    hipSetDevice(tr->device);
    TCHK(hipMemcpy(tr->feat, feat, sizeof(uint16_t) * 32 * n, hipMemcpyHostToDevice));
    TCHK(hipMemcpy(tr->resid, resid, sizeof(float) * n, hipMemcpyHostToDevice));
    const uint32_t nn = (uint32_t)n;
    TCHK(hipMemcpy(tr->count, &nn, 4, hipMemcpyHostToDevice));
    tr->n = n;
    return FF_OK;
}

int ff_trainer_train(ff_trainer *tr, int steps, uint32_t batch, float lr, float *last_loss)
{
    if (!tr || steps < 0 || batch == 0 || batch % T2_KB || batch > (1u << 22)) {
        if (tr) tr->err = "batch must be a positive multiple of 2048 (at most 4194304)";
        return FF_E_ARG;
    }
    if (tr->n == 0) {
        tr->err = "no training samples";
        return FF_E_STATE;
    }
    hipSetDevice(tr->device);
    for (int s = 0; s < steps; s++) {
        const bool last = s + 1 == steps;
        int rc = train_batch(tr, batch, nullptr, lr, true, last ? last_loss : nullptr);
        if (rc) return rc;
    }
    TCHK(hipStreamSynchronize(tr->stream));
    return FF_OK;
}

int ff_trainer_grad(ff_trainer *tr, const uint32_t *index, uint32_t n, float *grad, float *loss)
{
    if (!tr || !index || n == 0 || n % T2_KB || n > (1u << 22)) {
        if (tr) tr->err = "n must be a positive multiple of 2048";
        return FF_E_ARG;
    }
    for (uint32_t i = 0; i < n; i++)
        if (index[i] >= tr->n) {
            tr->err = "sample index out of range";
            return FF_E_ARG;
        }
    hipSetDevice(tr->device);
    int rc = ensure_batch(tr, n);
    if (rc) return rc;
    TCHK(hipMemcpyAsync(tr->index, index, sizeof(uint32_t) * n, hipMemcpyHostToDevice, tr->stream));
    rc = train_batch(tr, n, tr->index, 0.0f, false, loss);
    if (rc) return rc;
    if (grad) TCHK(hipMemcpy(grad, tr->grad, sizeof(float) * P_COUNT, hipMemcpyDeviceToHost));
    return FF_OK;
}

int ff_trainer_params(ff_trainer *tr, float *out, size_t n)
{
    if (!tr || !out || n != P_COUNT) return FF_E_ARG;
    hipSetDevice(tr->device);
    TCHK(hipStreamSynchronize(tr->stream));
    TCHK(hipMemcpy(out, tr->w.P, sizeof(float) * P_COUNT, hipMemcpyDeviceToHost));
    return FF_OK;
}

}  // extern "C"
