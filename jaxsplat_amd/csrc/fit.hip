// Device side of the image-fitting loop for gfx950: the training step of the
// reference's examples/single_image.py (render of activated parameters, L2
// loss, Adam) as three fused kernels around the render forward/backward.
//
//  * fit_activate_kernel: quats / |quats|, sigmoid(colors), sigmoid(opacities)
//    (single_image.py:143-166); means and scales are used raw;
//  * fit_l2_kernel + fit_l2_sum_kernel: v_img = 2 (img - target) / numel and
//    the loss sum (single_image.py:26-29: mean of squares), the workgroups'
//    partial sums added in workgroup order;
//  * fit_adam_kernel: per Gaussian, the activation VJP (sigmoid' = s (1 - s);
//    d(q/|q|) = (I - q^ q^T) / |q|) then Adam with bias correction
//    (optax.adam defaults b1 = 0.9, b2 = 0.999, eps = 1e-8 added to sqrt(v^));
//    the step counter lives on the device so a captured graph can be
//    replayed, and a non-zero skip word (the render's capacity-overflow flag)
//    turns the step into a no-op, so an overflowing render never updates the
//    parameters;
//  * fit_finish_kernel advances the step counter unless the step was skipped.
// All stores are plain vector stores and nothing is cleared by a memset, so
// the whole step may be captured.
#include "../../include/jaxsplat_amd.h"
#include "common.hpp"

namespace jsa {
namespace {

__device__ __forceinline__ float sigmoidf_(float x) { return 1.f / (1.f + __expf(-x)); }

__global__ __launch_bounds__(256) void fit_activate_kernel(
    int n, const float4* __restrict__ quats_raw, const float* __restrict__ colors_raw,
    const float* __restrict__ opac_raw, float4* __restrict__ quats, float* __restrict__ colors,
    float* __restrict__ opac) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 q = quats_raw[i];
    const float inv = rsqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    quats[i] = make_float4(q.x * inv, q.y * inv, q.z * inv, q.w * inv);
#pragma unroll
    for (int k = 0; k < 3; ++k) colors[3 * i + k] = sigmoidf_(colors_raw[3 * i + k]);
    opac[i] = sigmoidf_(opac_raw[i]);
}

// The loss's per-workgroup partial sums (one set per GPU, in the library):
// fit_l2_kernel writes them and fit_l2_sum_kernel adds them in workgroup
// order, so the loss does not depend on scheduling and no atomics on one
// address serialise the launch's end.  Calls on one GPU are issued in stream
// order by the callers (the fitting loop, the bench); calls that overlap would
// share the partials.
constexpr int kL2Blocks = 2048;
__device__ float l2_partial[kL2Blocks];

// v_img = scale (img - target) and the workgroup's partial loss sum; float4
// accesses when the three arrays are 16-byte aligned and n is a multiple of 4.
// Each thread takes kL2Unroll items kL2Unroll * 256 apart per pass and issues
// all their loads before the first store, so a pass keeps 2 kL2Unroll loads
// of every thread in flight (a 1080p image is one pass of the full grid).
constexpr int kL2Unroll = 4;
__global__ __launch_bounds__(256) void fit_l2_kernel(size_t n, const float* __restrict__ img,
                                                      const float* __restrict__ target,
                                                      float scale, float* __restrict__ v_img,
                                                      int vec4) {
    __shared__ float s_part[4];
    float acc = 0.f;
    const size_t stride = (size_t)gridDim.x * 256 * kL2Unroll;
    const size_t first = blockIdx.x * (size_t)(256 * kL2Unroll) + threadIdx.x;
    if (vec4) {
        const float4* a4 = reinterpret_cast<const float4*>(img);
        const float4* t4 = reinterpret_cast<const float4*>(target);
        float4* v4 = reinterpret_cast<float4*>(v_img);
        const size_t m = n / 4;
        for (size_t i0 = first; i0 < m; i0 += stride) {
            float4 a[kL2Unroll], t[kL2Unroll];
#pragma unroll
            for (int k = 0; k < kL2Unroll; ++k) {
                const size_t i = i0 + (size_t)k * 256;
                a[k] = i < m ? a4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
                t[k] = i < m ? t4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int k = 0; k < kL2Unroll; ++k) {
                const size_t i = i0 + (size_t)k * 256;
                const float4 d = make_float4(a[k].x - t[k].x, a[k].y - t[k].y, a[k].z - t[k].z,
                                             a[k].w - t[k].w);
                if (i < m) v4[i] = make_float4(scale * d.x, scale * d.y, scale * d.z, scale * d.w);
                acc += (d.x * d.x + d.y * d.y) + (d.z * d.z + d.w * d.w);
            }
        }
    } else {
        for (size_t i0 = first; i0 < n; i0 += stride) {
            float a[kL2Unroll], t[kL2Unroll];
#pragma unroll
            for (int k = 0; k < kL2Unroll; ++k) {
                const size_t i = i0 + (size_t)k * 256;
                a[k] = i < n ? img[i] : 0.f;
                t[k] = i < n ? target[i] : 0.f;
            }
#pragma unroll
            for (int k = 0; k < kL2Unroll; ++k) {
                const size_t i = i0 + (size_t)k * 256;
                const float d = a[k] - t[k];
                if (i < n) v_img[i] = scale * d;
                acc += d * d;
            }
        }
    }
    acc = wave_sum_to_lane63(acc);
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    if (lane == 63) s_part[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) l2_partial[blockIdx.x] = (s_part[0] + s_part[1]) + (s_part[2] + s_part[3]);
}

// loss_sum = the partials of the first `blocks` workgroups, in order (one
// workgroup of 256 threads).
__global__ __launch_bounds__(256) void fit_l2_sum_kernel(int blocks, float* __restrict__ loss_sum) {
    __shared__ float s_part[4];
    float acc = 0.f;
    for (int b = threadIdx.x; b < blocks; b += 256) acc += l2_partial[b];
    acc = wave_sum_to_lane63(acc);
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    if (lane == 63) s_part[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) *loss_sum = (s_part[0] + s_part[1]) + (s_part[2] + s_part[3]);
}

__device__ __forceinline__ float adam_one(float& p, float& m, float& v, float g, float lr,
                                          float b1, float b2, float eps, float c1, float c2) {
    m = b1 * m + (1.f - b1) * g;
    v = b2 * v + (1.f - b2) * g * g;
    const float mh = m / c1, vh = v / c2;
    p = p - lr * mh / (sqrtf(vh) + eps);
    return p;
}

// Parameter groups per Gaussian, in the order of the moment rows:
// means 3 | scales 3 | quats 4 | colors 3 | opacity 1  (14 floats).
__global__ __launch_bounds__(256) void fit_adam_kernel(
    int n, float* __restrict__ means, float* __restrict__ scales, float* __restrict__ quats_raw,
    float* __restrict__ colors_raw, float* __restrict__ opac_raw,
    const float* __restrict__ v_means, const float* __restrict__ v_scales,
    const float* __restrict__ v_quats, const float* __restrict__ v_colors,
    const float* __restrict__ v_opac, float* __restrict__ m1, float* __restrict__ m2,
    const int* __restrict__ step, const uint32_t* __restrict__ skip, float lr, float b1,
    float b2, float eps) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (skip && *skip) return;
    const float t = (float)(*step + 1);
    const float c1 = 1.f - powf(b1, t), c2 = 1.f - powf(b2, t);
    float g[14], p[14];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        p[k] = means[3 * i + k];
        g[k] = v_means[3 * i + k];
        p[3 + k] = scales[3 * i + k];
        g[3 + k] = v_scales[3 * i + k];
    }
    // quaternion normalisation VJP: (g - q^ (q^ . g)) / |q|
    {
        float q[4], gq[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            q[k] = quats_raw[4 * i + k];
            gq[k] = v_quats[4 * i + k];
        }
        const float nrm = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        const float inv = 1.f / nrm;
        float dot = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) dot += q[k] * inv * gq[k];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            p[6 + k] = q[k];
            g[6 + k] = (gq[k] - q[k] * inv * dot) * inv;
        }
    }
    // sigmoid VJP: g s (1 - s)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float x = colors_raw[3 * i + k];
        const float s = sigmoidf_(x);
        p[10 + k] = x;
        g[10 + k] = v_colors[3 * i + k] * s * (1.f - s);
    }
    {
        const float x = opac_raw[i];
        const float s = sigmoidf_(x);
        p[13] = x;
        g[13] = v_opac[i] * s * (1.f - s);
    }
    float mm[14], vv[14];
#pragma unroll
    for (int k = 0; k < 14; ++k) {
        mm[k] = m1[14 * (size_t)i + k];
        vv[k] = m2[14 * (size_t)i + k];
        adam_one(p[k], mm[k], vv[k], g[k], lr, b1, b2, eps, c1, c2);
        m1[14 * (size_t)i + k] = mm[k];
        m2[14 * (size_t)i + k] = vv[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        means[3 * i + k] = p[k];
        scales[3 * i + k] = p[3 + k];
        colors_raw[3 * i + k] = p[10 + k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) quats_raw[4 * i + k] = p[6 + k];
    opac_raw[i] = p[13];
}

__global__ void fit_finish_kernel(int* __restrict__ step, const uint32_t* __restrict__ skip) {
    if (!(skip && *skip)) *step += 1;
}

}  // namespace
}  // namespace jsa

extern "C" {

int jaxsplat_fit_activate(int64_t n, const float* quats_raw, const float* colors_raw,
                          const float* opacities_raw, float* quats, float* colors,
                          float* opacities, void* stream) {
    if (n < 0 || n > (1LL << 30)) return JAXSPLAT_ERR_INVALID;
    if (n == 0) return JAXSPLAT_OK;
    if (!quats_raw || !colors_raw || !opacities_raw || !quats || !colors || !opacities)
        return JAXSPLAT_ERR_INVALID;
    hipStream_t s = static_cast<hipStream_t>(stream);
    jsa::fit_activate_kernel<<<(int)((n + 255) / 256), 256, 0, s>>>(
        (int)n, reinterpret_cast<const float4*>(quats_raw), colors_raw, opacities_raw,
        reinterpret_cast<float4*>(quats), colors, opacities);
    return hipGetLastError() == hipSuccess ? JAXSPLAT_OK : JAXSPLAT_ERR_HIP;
}

int jaxsplat_fit_l2(int64_t n, const float* img, const float* target, float* v_img,
                    float* loss_sum, void* stream) {
    if (n <= 0 || !img || !target || !v_img || !loss_sum) return JAXSPLAT_ERR_INVALID;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const auto aligned = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
    const int vec4 = (n % 4 == 0 && aligned(img) && aligned(target) && aligned(v_img)) ? 1 : 0;
    const int64_t items = vec4 ? n / 4 : n;
    const int64_t per_block = 256 * jsa::kL2Unroll;
    const int64_t want = (items + per_block - 1) / per_block;
    const int blocks = (int)(want < jsa::kL2Blocks ? want : jsa::kL2Blocks);
    jsa::fit_l2_kernel<<<blocks, 256, 0, s>>>((size_t)n, img, target, 2.f / (float)n, v_img,
                                               vec4);
    if (hipGetLastError() != hipSuccess) return JAXSPLAT_ERR_HIP;
    jsa::fit_l2_sum_kernel<<<1, 256, 0, s>>>(blocks, loss_sum);
    return hipGetLastError() == hipSuccess ? JAXSPLAT_OK : JAXSPLAT_ERR_HIP;
}

int jaxsplat_fit_adam(int64_t n, float* means3d, float* scales, float* quats_raw,
                      float* colors_raw, float* opacities_raw, const float* v_means3d,
                      const float* v_scales, const float* v_quats, const float* v_colors,
                      const float* v_opacities, float* moment1, float* moment2, int32_t* step,
                      const uint32_t* skip, float lr, float beta1, float beta2, float eps,
                      void* stream) {
    if (n < 0 || n > (1LL << 30)) return JAXSPLAT_ERR_INVALID;
    if (n == 0) return JAXSPLAT_OK;
    if (!means3d || !scales || !quats_raw || !colors_raw || !opacities_raw || !v_means3d ||
        !v_scales || !v_quats || !v_colors || !v_opacities || !moment1 || !moment2 || !step)
        return JAXSPLAT_ERR_INVALID;
    hipStream_t s = static_cast<hipStream_t>(stream);
    jsa::fit_adam_kernel<<<(int)((n + 255) / 256), 256, 0, s>>>(
        (int)n, means3d, scales, quats_raw, colors_raw, opacities_raw, v_means3d, v_scales,
        v_quats, v_colors, v_opacities, moment1, moment2, step, skip, lr, beta1, beta2, eps);
    jsa::fit_finish_kernel<<<1, 1, 0, s>>>(step, skip);
    return hipGetLastError() == hipSuccess ? JAXSPLAT_OK : JAXSPLAT_ERR_HIP;
}

}  // extern "C"
