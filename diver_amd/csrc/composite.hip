// Alpha compositing of dense per-ray samples (model/diver.py:183-207) and its adjoint.
//   alpha_i = 1 - exp(-relu(sigma_i) * mask_i);  T_i = prod_{j<i} (1 - alpha_j)
//   weight_i = alpha_i T_i;  rgb = sum_i weight_i c_i (+ 1 - sum_i weight_i with a white background)
// One ray per lane.  The adjoint uses the suffix recursion
//   U_k = e_{k+1} alpha_{k+1} + (1 - alpha_{k+1}) U_{k+1},   dL/dalpha_k = T_k (e_k - U_k),
// with e_i = dL/dweight_i, which never divides by (1 - alpha).
#include "common.h"
#include "kernels.h"

namespace diver {

__global__ void __launch_bounds__(256) composite_kernel(int64_t batch, int m,
                                                        const float* __restrict__ sigma,
                                                        const float* __restrict__ color,
                                                        const uint8_t* __restrict__ mask,
                                                        int white_back, float* __restrict__ rgb,
                                                        float* __restrict__ weight) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    float T = 1.0f, r = 0.0f, g = 0.0f, bl = 0.0f, wsum = 0.0f;
    for (int i = 0; i < m; ++i) {
        int64_t k = b * m + i;
        float s = fmaxf(sigma[k], 0.0f);
        float a = 1.0f - expf(-(mask[k] ? s : 0.0f));
        float w = a * T;
        weight[k] = w;
        wsum += w;
        r += w * color[3 * k + 0];
        g += w * color[3 * k + 1];
        bl += w * color[3 * k + 2];
        T = T * (1.0f - a);
    }
    if (white_back) {
        r += 1.0f - wsum;
        g += 1.0f - wsum;
        bl += 1.0f - wsum;
    }
    rgb[3 * b + 0] = r;
    rgb[3 * b + 1] = g;
    rgb[3 * b + 2] = bl;
}

__global__ void __launch_bounds__(256) composite_backward_kernel(
    int64_t batch, int m, const float* __restrict__ sigma, const float* __restrict__ color,
    const uint8_t* __restrict__ mask, int white_back, const float* __restrict__ grad_rgb,
    const float* __restrict__ grad_weight, float* __restrict__ grad_sigma,
    float* __restrict__ grad_color) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const float gr = grad_rgb[3 * b], gg = grad_rgb[3 * b + 1], gb = grad_rgb[3 * b + 2];
    const float gwb = white_back ? (gr + gg + gb) : 0.0f;
    // forward pass for T_i (kept in the output buffer to avoid a second array)
    float T = 1.0f;
    for (int i = 0; i < m; ++i) {
        int64_t k = b * m + i;
        float s = fmaxf(sigma[k], 0.0f);
        float a = 1.0f - expf(-(mask[k] ? s : 0.0f));
        float w = a * T;
        grad_color[3 * k + 0] = gr * w;
        grad_color[3 * k + 1] = gg * w;
        grad_color[3 * k + 2] = gb * w;
        grad_sigma[k] = T;
        T = T * (1.0f - a);
    }
    float U = 0.0f;
    for (int i = m - 1; i >= 0; --i) {
        int64_t k = b * m + i;
        float s_raw = sigma[k];
        bool on = mask[k] && s_raw > 0.0f;
        float s = fmaxf(s_raw, 0.0f);
        float ea = expf(-(mask[k] ? s : 0.0f));  // 1 - alpha
        float a = 1.0f - ea;
        float e = gr * color[3 * k] + gg * color[3 * k + 1] + gb * color[3 * k + 2] - gwb;
        if (grad_weight) e += grad_weight[k];
        float Ti = grad_sigma[k];
        float dalpha = Ti * (e - U);
        grad_sigma[k] = on ? dalpha * ea : 0.0f;
        U = e * a + ea * U;
    }
}

hipError_t launch_composite(int64_t batch, int m, const float* sigma, const float* color,
                            const uint8_t* mask, int white_back, float* rgb, float* weight,
                            hipStream_t s) {
    if (batch == 0) return hipSuccess;
    hipLaunchKernelGGL(composite_kernel, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s,
                       batch, m, sigma, color, mask, white_back, rgb, weight);
    return hipGetLastError();
}

hipError_t launch_composite_backward(int64_t batch, int m, const float* sigma, const float* color,
                                     const uint8_t* mask, int white_back, const float* grad_rgb,
                                     const float* grad_weight, float* grad_sigma,
                                     float* grad_color, hipStream_t s) {
    if (batch == 0) return hipSuccess;
    hipLaunchKernelGGL(composite_backward_kernel, dim3((unsigned)((batch + 255) / 256)), dim3(256),
                       0, s, batch, m, sigma, color, mask, white_back, grad_rgb, grad_weight,
                       grad_sigma, grad_color);
    return hipGetLastError();
}

}  // namespace diver
