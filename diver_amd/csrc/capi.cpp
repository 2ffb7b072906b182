// C-ABI of libdiver: argument validation, decoder handles, dispatch to the HIP launchers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/diver.h"
#include "kernels.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_status(hipError_t e, const char* what) {
    if (e == hipSuccess) return DIVER_OK;
    return fail(DIVER_EHIP, "%s: %s", what, hipGetErrorString(e));
}

#define REQUIRE(cond, ...)                               \
    do {                                                 \
        if (!(cond)) return fail(DIVER_EINVAL, __VA_ARGS__); \
    } while (0)

hipStream_t S(diver_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

int variant_of(const diver_decoder_desc& d) {
    if (d.n_pre == 0 && !d.pre_identity && d.density_act == 1 && d.n_post == 1) {
        if (d.channels == 32 && d.hidden == 32) return 0;
        if (d.channels == 32 && d.hidden == 64) return 1;
        if (d.channels == 64 && d.hidden == 64) return 2;
    }
    if (d.density_act == 0 && d.n_post == 0 && d.channels == 32) {
        if (d.hidden == 32 && d.n_pre == 2 && !d.pre_identity) return 3;
        if (d.hidden == 64 && d.n_pre == 1 && !d.pre_identity) return 4;
        if (d.hidden == 32 && d.n_pre == 2 && d.pre_identity) return 5;
    }
    return -1;
}

uint16_t to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

uint16_t to_f16(float f) {
    _Float16 h = (_Float16)f;
    uint16_t r;
    std::memcpy(&r, &h, 2);
    return r;
}

}  // namespace

namespace diver {
// Sets the calling thread's diver_last_error() text (used by the render server's entry points).
int set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
}  // namespace diver

struct diver_decoder {
    diver_decoder_desc desc;
    int variant;
    int device;
    int64_t n_params;
    bool density_ch0 = false;      // sigma = act(F[0]): w_s = e_0, b_s = 0, no pre-layers
    bool packs_valid = true;       // matrix-core packs match W (false after a device update)
    float* W = nullptr;            // canonical fp32
    uint32_t* pack_bf16 = nullptr; // matrix-core A operands (32x32x16 layout, round kernel)
    uint32_t* pack_f16 = nullptr;
    uint32_t* pack16_bf16 = nullptr;  // matrix-core A operands (16x16x32 layout, tile kernel)
    uint32_t* pack16_f16 = nullptr;
    float* pack_f32 = nullptr;     // FP32 matrix-core A operands (repacked on every update)
    float* pack_pre = nullptr;     // FP32 matrix-core A operands of the pre-layers (repacked on every update)
    float* pack_adj = nullptr;     // FP32 transposed operands of the decoder adjoint (repacked on every update)
};

extern "C" {

const char* diver_version(void) { return "libdiver 0.1 gfx950"; }

const char* diver_last_error(void) { return g_err.c_str(); }

int diver_ray_voxel_intersect(int64_t batch, int max_n, float xyzmin, float xyzmax,
                              float voxel_num, float voxel_size, const float* o, const float* v,
                              float* isect, int32_t* num, float* tns, diver_stream_t stream) {
    (void)xyzmax;
    REQUIRE(batch >= 0, "batch < 0");
    REQUIRE(batch == 0 || (o && v && num), "null ray or count pointer");
    REQUIRE(max_n >= 0, "max_n < 0");
    REQUIRE(voxel_num > 0 && voxel_size > 0, "voxel_num and voxel_size must be positive");
    return hip_status(diver::launch_ray_voxel_intersect(batch, max_n, xyzmin, voxel_num,
                                                        voxel_size, o, v, isect, num, tns,
                                                        S(stream)),
                      "ray_voxel_intersect");
}

int diver_masked_intersect(int64_t batch, int max_n, float xyzmin, float xyzmax, float voxel_num,
                           float voxel_size, float mask_scale, const float* o, const float* v,
                           const uint8_t* mask, float* isect, int32_t* num, float* tns,
                           diver_stream_t stream) {
    (void)xyzmax;
    REQUIRE(batch >= 0, "batch < 0");
    REQUIRE(batch == 0 || (o && v && num && mask), "null ray, mask or count pointer");
    REQUIRE(max_n >= 0, "max_n < 0");
    REQUIRE(voxel_num > 0 && voxel_size > 0 && mask_scale > 0,
            "voxel_num, voxel_size and mask_scale must be positive");
    return hip_status(diver::launch_masked_intersect(batch, max_n, xyzmin, voxel_num, voxel_size,
                                                     mask_scale, o, v, mask, isect, num, tns,
                                                     S(stream)),
                      "masked_intersect");
}

int diver_aabb_intersect(int64_t batch, float* coords, const float* v, const float* center,
                         uint8_t* finish, float* rgba, const uint8_t* octrees, int voxel_num,
                         diver_stream_t stream) {
    REQUIRE(batch >= 0, "batch < 0");
    REQUIRE(voxel_num >= 64 && voxel_num % 64 == 0,
            "voxel_num must be a positive multiple of 64 (got %d)", voxel_num);
    REQUIRE(batch == 0 || (coords && v && center && finish && rgba && octrees), "null pointer");
    return hip_status(diver::launch_aabb_intersect(batch, coords, v, center, finish, rgba, octrees,
                                                   voxel_num, S(stream)),
                      "aabb_intersect");
}

int diver_ray_march(int64_t batch, float* coords, const float* v, const uint8_t* mask,
                    const uint8_t* finish, int voxel_num, int64_t thread_stride,
                    diver_stream_t stream) {
    REQUIRE(batch >= 0, "batch < 0");
    REQUIRE(voxel_num > 0, "voxel_num must be positive");
    REQUIRE(batch == 0 || (coords && v && mask && finish), "null pointer");
    if (thread_stride <= 0) thread_stride = batch > 0 ? batch : 1;
    return hip_status(diver::launch_ray_march(batch, coords, v, mask, finish, voxel_num,
                                              thread_stride, S(stream)),
                      "ray_march");
}

int diver_integrate(const float* voxels, int n1, int channels, int64_t k, const float* coord_in,
                    const float* coord_out, float* feature, diver_stream_t stream) {
    REQUIRE(k >= 0 && n1 > 0, "bad sizes");
    REQUIRE(channels > 0 && channels % 4 == 0, "channels must be a positive multiple of 4");
    REQUIRE(k == 0 || (voxels && coord_in && coord_out && feature), "null pointer");
    return hip_status(diver::launch_integrate(voxels, n1, channels, k, coord_in, coord_out,
                                              feature, S(stream)),
                      "integrate");
}

int diver_integrate_backward(int n1, int channels, int64_t k, const float* coord_in,
                             const float* coord_out, const float* grad_feature,
                             float* grad_voxels, diver_stream_t stream) {
    REQUIRE(k >= 0 && n1 > 0, "bad sizes");
    REQUIRE(channels > 0 && channels % 4 == 0, "channels must be a positive multiple of 4");
    REQUIRE(k == 0 || (grad_voxels && coord_in && coord_out && grad_feature), "null pointer");
    return hip_status(diver::launch_integrate_backward(n1, channels, k, coord_in, coord_out,
                                                       grad_feature, grad_voxels, S(stream)),
                      "integrate_backward");
}

int diver_corner_keys(int n1, int64_t k, const float* coord_in, const float* coord_out,
                      int64_t* keys, diver_stream_t stream) {
    REQUIRE(k >= 0 && n1 > 1, "bad sizes");
    REQUIRE(k == 0 || (coord_in && coord_out && keys), "null pointer");
    return hip_status(diver::launch_corner_keys(n1, k, coord_in, coord_out, keys, S(stream)),
                      "corner_keys");
}

int diver_integrate_table(const float* table, int channels, int64_t k, const int64_t* corner_row,
                          const float* coord_in, const float* coord_out, float* feature,
                          diver_stream_t stream) {
    REQUIRE(k >= 0, "bad sizes");
    REQUIRE(channels > 0 && channels % 4 == 0, "channels must be a positive multiple of 4");
    REQUIRE(k == 0 || (table && corner_row && coord_in && coord_out && feature), "null pointer");
    return hip_status(diver::launch_integrate_table(table, channels, k, corner_row, coord_in,
                                                    coord_out, feature, S(stream)),
                      "integrate_table");
}

int diver_integrate_table_backward(int channels, int64_t k, const int64_t* corner_row,
                                   const float* coord_in, const float* coord_out,
                                   const float* grad_feature, float* grad_table,
                                   diver_stream_t stream) {
    REQUIRE(k >= 0, "bad sizes");
    REQUIRE(channels > 0 && channels % 4 == 0, "channels must be a positive multiple of 4");
    REQUIRE(k == 0 || (corner_row && coord_in && coord_out && grad_feature && grad_table),
            "null pointer");
    return hip_status(diver::launch_integrate_table_backward(channels, k, corner_row, coord_in,
                                                             coord_out, grad_feature, grad_table,
                                                             S(stream)),
                      "integrate_table_backward");
}

int diver_composite(int64_t batch, int m, const float* sigma, const float* color,
                    const uint8_t* mask, int white_back, float* rgb, float* weight,
                    diver_stream_t stream) {
    REQUIRE(batch >= 0 && m >= 0, "bad sizes");
    REQUIRE(batch == 0 || (rgb && (m == 0 || (sigma && color && mask && weight))), "null pointer");
    return hip_status(
        diver::launch_composite(batch, m, sigma, color, mask, white_back, rgb, weight, S(stream)),
        "composite");
}

int diver_composite_backward(int64_t batch, int m, const float* sigma, const float* color,
                             const uint8_t* mask, int white_back, const float* grad_rgb,
                             const float* grad_weight, float* grad_sigma, float* grad_color,
                             diver_stream_t stream) {
    REQUIRE(batch >= 0 && m >= 0, "bad sizes");
    REQUIRE(batch == 0 || m == 0 || (sigma && color && mask && grad_rgb && grad_sigma && grad_color),
            "null pointer");
    return hip_status(diver::launch_composite_backward(batch, m, sigma, color, mask, white_back,
                                                       grad_rgb, grad_weight, grad_sigma,
                                                       grad_color, S(stream)),
                      "composite_backward");
}

int64_t diver_decoder_param_count(const diver_decoder_desc* d) {
    if (!d) return -1;
    const int64_t C = d->channels, H = d->hidden;
    const int64_t Z = d->n_pre ? H : C;
    int64_t n = 0;
    for (int i = 0; i < d->n_pre; ++i) n += (i == 0) ? (d->pre_identity ? H : H * C + H) : H * H + H;
    n += Z + 1 + H * Z + H * 27 + H;
    n += (int64_t)d->n_post * (H * H + H);
    n += 3 * H + 3;
    return n;
}

// Matrix-core A operands of one layer W[M][K] (row-major): per (32-row tile, 16-wide k chunk,
// lane) 8 consecutive k values of row 32*mt + lane%32 starting at 16*kc + 8*(lane/32).
static void pack_layer(const float* W, int M, int K, bool bf16, std::vector<uint32_t>& out) {
    const int MT = (M + 31) / 32, KC = K / 16;
    for (int mt = 0; mt < MT; ++mt)
        for (int kc = 0; kc < KC; ++kc)
            for (int lane = 0; lane < 64; ++lane) {
                int row = 32 * mt + (lane & 31);
                int k0 = 16 * kc + 8 * (lane >> 5);
                for (int q = 0; q < 4; ++q) {
                    float a = row < M ? W[(int64_t)row * K + k0 + 2 * q] : 0.0f;
                    float b = row < M ? W[(int64_t)row * K + k0 + 2 * q + 1] : 0.0f;
                    uint16_t ha = bf16 ? to_bf16(a) : to_f16(a);
                    uint16_t hb = bf16 ? to_bf16(b) : to_f16(b);
                    out.push_back((uint32_t)ha | ((uint32_t)hb << 16));
                }
            }
}

// Matrix-core A operands of one layer W[M][K] for v_mfma_f32_16x16x32: per (16-row tile mt,
// 32-wide k step kk, lane) the 8 k values of row 16*mt + lane%16 that lane group g = lane/16
// supplies, in the order the tile kernel's B operands arrive in: 32*kk + 4*g + j for j < 4 and
// 32*kk + 16 + 4*g + (j - 4) for j >= 4.  For the head layer these are the channels its
// 4-lanes-per-row gather sums in a lane, for the hidden layers the units a lane holds in D tiles
// 2*kk and 2*kk + 1.
static void pack16_layer(const float* W, int M, int K, bool bf16, std::vector<uint32_t>& out) {
    const int MT = (M + 15) / 16, KS = K / 32;
    for (int mt = 0; mt < MT; ++mt)
        for (int kk = 0; kk < KS; ++kk)
            for (int lane = 0; lane < 64; ++lane) {
                const int row = 16 * mt + (lane & 15), g = lane >> 4;
                auto kof = [&](int j) {
                    return j < 4 ? 32 * kk + 4 * g + j : 32 * kk + 16 + 4 * g + (j - 4);
                };
                for (int q = 0; q < 4; ++q) {
                    float a = row < M ? W[(int64_t)row * K + kof(2 * q)] : 0.0f;
                    float b = row < M ? W[(int64_t)row * K + kof(2 * q + 1)] : 0.0f;
                    uint16_t ha = bf16 ? to_bf16(a) : to_f16(a);
                    uint16_t hb = bf16 ? to_bf16(b) : to_f16(b);
                    out.push_back((uint32_t)ha | ((uint32_t)hb << 16));
                }
            }
}

int diver_decoder_create(const diver_decoder_desc* desc, const float* params_host,
                         int64_t n_params, diver_decoder** out) {
    REQUIRE(desc && params_host && out, "null argument");
    *out = nullptr;
    int variant = variant_of(*desc);
    if (variant < 0)
        return fail(DIVER_EUNSUPPORTED,
                    "decoder (channels=%d hidden=%d n_pre=%d pre_identity=%d act=%d n_post=%d) is "
                    "not compiled into this library",
                    desc->channels, desc->hidden, desc->n_pre, desc->pre_identity,
                    desc->density_act, desc->n_post);
    int64_t need = diver_decoder_param_count(desc);
    REQUIRE(n_params == need, "decoder expects %lld parameters, got %lld", (long long)need,
            (long long)n_params);
    for (int64_t i = 0; i < n_params; ++i)
        REQUIRE(std::isfinite(params_host[i]), "non-finite decoder parameter at %lld", (long long)i);

    diver_decoder* d = new diver_decoder();
    d->desc = *desc;
    d->variant = variant;
    d->n_params = n_params;
    hipGetDevice(&d->device);
    auto cleanup = [&](int code, const char* what) {
        if (d->W) (void)hipFree(d->W);
        if (d->pack_bf16) (void)hipFree(d->pack_bf16);
        if (d->pack_f16) (void)hipFree(d->pack_f16);
        if (d->pack16_bf16) (void)hipFree(d->pack16_bf16);
        if (d->pack16_f16) (void)hipFree(d->pack16_f16);
        if (d->pack_f32) (void)hipFree(d->pack_f32);
        if (d->pack_pre) (void)hipFree(d->pack_pre);
        if (d->pack_adj) (void)hipFree(d->pack_adj);
        delete d;
        return fail(code, "%s", what);
    };
    if (hipMalloc(&d->W, sizeof(float) * n_params) != hipSuccess)
        return cleanup(DIVER_ENOMEM, "decoder parameter allocation failed");
    if (hipMemcpy(d->W, params_host, sizeof(float) * n_params, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(DIVER_EHIP, "decoder parameter upload failed");

    if (desc->n_pre == 0) {
        const int C = desc->channels, H = desc->hidden;
        bool ch0 = params_host[0] == 1.0f && params_host[C] == 0.0f;
        for (int i = 1; i < C; ++i) ch0 = ch0 && params_host[i] == 0.0f;
        d->density_ch0 = ch0;
        const int64_t wc = (int64_t)C + 1;  // w_s[Z], b_s then W_c
        const int64_t post0 = wc + (int64_t)H * C + (int64_t)H * 27 + H;
        const int64_t outo = post0 + (int64_t)desc->n_post * (H * H + H);
        for (int prec = 0; prec < 2; ++prec) {
            std::vector<uint32_t> pk;
            pack_layer(params_host + wc, H, C, prec == 0, pk);
            for (int j = 0; j < desc->n_post; ++j)
                pack_layer(params_host + post0 + (int64_t)j * (H * H + H), H, H, prec == 0, pk);
            pack_layer(params_host + outo, 3, H, prec == 0, pk);
            uint32_t** dst = prec == 0 ? &d->pack_bf16 : &d->pack_f16;
            if (hipMalloc(dst, pk.size() * 4) != hipSuccess)
                return cleanup(DIVER_ENOMEM, "decoder pack allocation failed");
            if (hipMemcpy(*dst, pk.data(), pk.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
                return cleanup(DIVER_EHIP, "decoder pack upload failed");
            if (C % 32 == 0 && H % 32 == 0) {
                std::vector<uint32_t> p16;
                pack16_layer(params_host + wc, H, C, prec == 0, p16);
                for (int j = 0; j < desc->n_post; ++j)
                    pack16_layer(params_host + post0 + (int64_t)j * (H * H + H), H, H, prec == 0,
                                 p16);
                pack16_layer(params_host + outo, 3, H, prec == 0, p16);
                uint32_t** dst16 = prec == 0 ? &d->pack16_bf16 : &d->pack16_f16;
                if (hipMalloc(dst16, p16.size() * 4) != hipSuccess)
                    return cleanup(DIVER_ENOMEM, "decoder pack allocation failed");
                if (hipMemcpy(*dst16, p16.data(), p16.size() * 4, hipMemcpyHostToDevice) !=
                    hipSuccess)
                    return cleanup(DIVER_EHIP, "decoder pack upload failed");
            }
        }
        const int64_t nf = diver::pack_f32_size(C, H, desc->n_post);
        if (hipMalloc(&d->pack_f32, sizeof(float) * nf) != hipSuccess)
            return cleanup(DIVER_ENOMEM, "decoder pack allocation failed");
        if (diver::launch_pack_f32(C, H, desc->n_post, d->W, d->pack_f32, nullptr) != hipSuccess ||
            hipStreamSynchronize(nullptr) != hipSuccess)
            return cleanup(DIVER_EHIP, "decoder FP32 pack failed");
        if (C % 32 == 0 && H % 32 == 0 && desc->n_post == 1) {
            const int64_t na = diver::pack_adj_size(C, H, desc->n_post);
            if (hipMalloc(&d->pack_adj, sizeof(float) * na) != hipSuccess)
                return cleanup(DIVER_ENOMEM, "decoder pack allocation failed");
            if (diver::launch_pack_adj(C, H, desc->n_post, d->W, d->pack_adj, nullptr) != hipSuccess ||
                hipStreamSynchronize(nullptr) != hipSuccess)
                return cleanup(DIVER_EHIP, "decoder adjoint pack failed");
        }
    } else if (!desc->pre_identity && desc->channels % 16 == 0 && desc->hidden % 32 == 0) {
        const int64_t nf = diver::pack_pre_f32_size(desc->channels, desc->hidden, desc->n_pre);
        if (hipMalloc(&d->pack_pre, sizeof(float) * nf) != hipSuccess)
            return cleanup(DIVER_ENOMEM, "decoder pack allocation failed");
        if (diver::launch_pack_pre_f32(desc->channels, desc->hidden, desc->n_pre, d->W, d->pack_pre,
                                       nullptr) != hipSuccess ||
            hipStreamSynchronize(nullptr) != hipSuccess)
            return cleanup(DIVER_EHIP, "decoder FP32 pre-layer pack failed");
    }
    *out = d;
    return DIVER_OK;
}

int diver_decoder_update(diver_decoder* d, const float* params, int64_t n_params,
                         diver_stream_t stream) {
    REQUIRE(d && params, "null argument");
    REQUIRE(n_params == d->n_params, "decoder expects %lld parameters, got %lld",
            (long long)d->n_params, (long long)n_params);
    return diver_decoder_update_range(d, params, 0, n_params, stream);
}

int diver_decoder_update_range(diver_decoder* d, const float* params, int64_t first,
                               int64_t count, diver_stream_t stream) {
    REQUIRE(d && params, "null argument");
    REQUIRE(first >= 0 && count >= 0 && first <= d->n_params && count <= d->n_params - first,
            "parameter range [%lld, %lld) outside the decoder's %lld parameters", (long long)first,
            (long long)(first + count), (long long)d->n_params);
    d->packs_valid = false;
    // the channel-0 density shortcut holds while w_s / b_s (the first channels + 1 parameters of a
    // decoder without pre-layers) keep the values it was detected on
    if (first < (int64_t)d->desc.channels + 1) d->density_ch0 = false;
    if (count == 0) return DIVER_OK;
    int st = hip_status(hipMemcpyAsync(d->W + first, params, sizeof(float) * count,
                                       hipMemcpyDeviceToDevice, S(stream)),
                        "decoder_update");
    if (st != DIVER_OK) return st;
    // the FP32 operands follow W on the same stream
    if (d->pack_pre)
        return hip_status(diver::launch_pack_pre_f32(d->desc.channels, d->desc.hidden, d->desc.n_pre,
                                                     d->W, d->pack_pre, S(stream)),
                          "decoder_update (FP32 pre-layer pack)");
    if (!d->pack_f32) return st;
    st = hip_status(diver::launch_pack_f32(d->desc.channels, d->desc.hidden, d->desc.n_post, d->W,
                                           d->pack_f32, S(stream)),
                    "decoder_update (FP32 pack)");
    if (st != DIVER_OK || !d->pack_adj) return st;
    return hip_status(diver::launch_pack_adj(d->desc.channels, d->desc.hidden, d->desc.n_post, d->W,
                                             d->pack_adj, S(stream)),
                      "decoder_update (adjoint pack)");
}

int diver_decoder_destroy(diver_decoder* d) {
    if (!d) return DIVER_OK;
    int dev = 0;
    hipGetDevice(&dev);
    hipSetDevice(d->device);
    if (d->W) (void)hipFree(d->W);
    if (d->pack_bf16) (void)hipFree(d->pack_bf16);
    if (d->pack_f16) (void)hipFree(d->pack_f16);
    if (d->pack16_bf16) (void)hipFree(d->pack16_bf16);
    if (d->pack16_f16) (void)hipFree(d->pack16_f16);
    if (d->pack_f32) (void)hipFree(d->pack_f32);
    if (d->pack_pre) (void)hipFree(d->pack_pre);
    if (d->pack_adj) (void)hipFree(d->pack_adj);
    hipSetDevice(dev);
    delete d;
    return DIVER_OK;
}

int diver_mlp_eval(const diver_decoder* dec, int64_t batch, const float* coord,
                   const float* features, const int32_t* voxel_map, int map_size, const float* v,
                   float* rgba, uint8_t* finish, float alpha_skip, float t_stop,
                   diver_stream_t stream) {
    REQUIRE(dec, "null decoder");
    REQUIRE(batch >= 0 && map_size > 0, "bad sizes");
    REQUIRE(batch == 0 || (coord && features && voxel_map && v && rgba && finish), "null pointer");
    return hip_status(diver::launch_mlp_eval(dec->variant, batch, coord, features, voxel_map,
                                             map_size, v, rgba, finish, dec->W, alpha_skip, t_stop,
                                             S(stream)),
                      "mlp_eval");
}

namespace {
// The tile kernel indexes features with 32-bit offsets and keeps the block bits in LDS.
bool tiles_ok(const diver_decoder* dec, const diver::RenderArgs& a) {
    const double corners = (double)(a.n + 1) * (a.n + 1) * (a.n + 1) * a.channels;
    return dec->desc.n_pre == 0 && dec->variant >= 0 && dec->variant <= 2 && a.blocks && a.bricks &&
           corners < 4294967296.0 && diver::tile_fits(a.n);
}
}  // namespace

// Decoders without pre-layers use the tile kernel (per-lane walk + density, colour in 64-row
// matrix-core batches) when the grid's block and brick bits are available; everything else uses
// the round kernel.
static int dispatch_render(const diver_decoder* dec, const diver::RenderArgs& a, int precision,
                           hipStream_t s, const char* what) {
    if (!tiles_ok(dec, a))
        return hip_status(diver::launch_render(dec->variant, a, precision, s), what);
    diver::BatchArgs p{};
    p.g = a;
    p.n_req = 1;
    p.req[0].cam = a.cam;
    p.req[0].ray_start = 0;
    p.req[0].n_rays = a.n_rays;
    p.req[0].o = a.o;
    p.req[0].d = a.d;
    p.req[0].rgba = a.rgba;
    p.req[0].count = a.count;
    p.total = a.n_rays;
    return hip_status(diver::launch_render_tiles(dec->variant, p, precision, dec->density_ch0, s),
                      what);
}

int diver_walk_rays(const diver_grid* g, int traversal, int64_t n_rays, const float* o,
                    const float* d, int max_n, float* isect, int32_t* num,
                    diver_stream_t stream) {
    REQUIRE(g && g->mask && n_rays >= 0 && max_n >= 0, "bad arguments");
    REQUIRE(n_rays == 0 || (o && d && num && (max_n == 0 || isect)), "null pointer");
    REQUIRE(traversal == DIVER_WALK_OFFLINE || traversal == DIVER_WALK_REALTIME,
            "bad traversal %d", traversal);
    if (traversal == DIVER_WALK_REALTIME)
        REQUIRE(g->octrees && g->n % 64 == 0 && g->n >= 64 && g->n <= 4096,
                "the real-time walk needs the pyramid and n a multiple of 64 in [64, 4096]");
    REQUIRE(g->blocks && g->bricks && g->n % 8 == 0,
            "the walk needs the block and brick bits (n % 8 == 0)");
    diver::RenderArgs a{};
    a.mask = g->mask;
    a.octrees = g->octrees;
    a.blocks = g->blocks;
    a.bricks = reinterpret_cast<const unsigned long long*>(g->bricks);
    if (g->tail) REQUIRE(g->n % 8 == 0, "tail bits need n to be a multiple of 8");
    a.tail = g->tail;
    a.traversal = traversal;
    a.n = g->n;
    a.xyzmin = g->xyzmin;
    a.voxel_size = g->voxel_size;
    a.n_rays = n_rays;
    a.o = o;
    a.d = d;
    return hip_status(diver::launch_walk(a, max_n, isect, num, S(stream)), "walk_rays");
}

int diver_tail_bits(const uint32_t* blocks, int n, uint32_t* tail, diver_stream_t stream) {
    REQUIRE(blocks && tail, "null pointer");
    REQUIRE(n >= 8 && n % 8 == 0 && n <= 4096, "tail bits need n a multiple of 8 in [8, 4096]");
    return hip_status(diver::launch_tail_bits(blocks, n, tail, S(stream)), "tail_bits");
}

int diver_brick_bits(const uint8_t* mask, int n, uint64_t* bricks, diver_stream_t stream) {
    REQUIRE(mask && bricks, "null pointer");
    REQUIRE(n >= 4 && n % 4 == 0, "n must be a positive multiple of 4");
    return hip_status(diver::launch_brick_bits(mask, n, reinterpret_cast<unsigned long long*>(bricks),
                                               S(stream)),
                      "brick_bits");
}

int diver_adam_step(float* param, float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                    double beta1, double beta2, float eps, float weight_decay, float step_size,
                    float bc2_sqrt, int clear_grad, diver_stream_t stream) {
    REQUIRE(n >= 0, "n < 0");
    if (n == 0) return DIVER_OK;
    REQUIRE(param && grad && exp_avg && exp_avg_sq, "null pointer");
    const uintptr_t al = reinterpret_cast<uintptr_t>(param) | reinterpret_cast<uintptr_t>(grad) |
                         reinterpret_cast<uintptr_t>(exp_avg) |
                         reinterpret_cast<uintptr_t>(exp_avg_sq);
    REQUIRE(al % 4 == 0, "parameter, gradient and moment arrays must be float-aligned");
    REQUIRE(beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0, "betas outside [0, 1)");
    REQUIRE(eps >= 0.0f && bc2_sqrt > 0.0f && std::isfinite(step_size), "bad Adam scalars");
    return hip_status(diver::launch_adam(param, grad, exp_avg, exp_avg_sq, n, beta1, beta2, eps,
                                         weight_decay, step_size, bc2_sqrt, clear_grad != 0,
                                         S(stream)),
                      "adam_step");
}

int diver_adam_step_rows(float* param, float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                         const int64_t* rows, int64_t n_rows, int row_len, double beta1,
                         double beta2, float eps, float weight_decay, float step_size,
                         float bc2_sqrt, int clear_grad, diver_stream_t stream) {
    REQUIRE(n >= 0 && n_rows >= 0, "negative size");
    if (n_rows == 0) return DIVER_OK;
    REQUIRE(param && grad && exp_avg && exp_avg_sq && rows, "null pointer");
    REQUIRE(row_len > 0 && row_len % 4 == 0 && n % row_len == 0,
            "row_len must be a positive multiple of 4 dividing n");
    const uintptr_t al = reinterpret_cast<uintptr_t>(param) | reinterpret_cast<uintptr_t>(grad) |
                         reinterpret_cast<uintptr_t>(exp_avg) |
                         reinterpret_cast<uintptr_t>(exp_avg_sq);
    REQUIRE(al % 16 == 0, "parameter, gradient and moment arrays must be 16-byte aligned");
    REQUIRE(weight_decay == 0.0f,
            "a row-restricted step equals the full one only without weight decay");
    REQUIRE(beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0, "betas outside [0, 1)");
    REQUIRE(eps >= 0.0f && bc2_sqrt > 0.0f && std::isfinite(step_size), "bad Adam scalars");
    return hip_status(diver::launch_adam_rows(param, grad, exp_avg, exp_avg_sq, n, rows, n_rows,
                                              row_len, beta1, beta2, eps, step_size, bc2_sqrt,
                                              clear_grad != 0, S(stream)),
                      "adam_step_rows");
}

int diver_block_bits(const uint8_t* mask, int n, uint32_t* bits, diver_stream_t stream) {
    REQUIRE(mask && bits, "null pointer");
    REQUIRE(n >= 8 && n % 8 == 0, "n must be a positive multiple of 8");
    return hip_status(diver::launch_block_bits(mask, n, bits, S(stream)), "block_bits");
}

static int render_common(const diver_grid* g, const diver_decoder* dec,
                         const diver_render_opts* opts, diver::RenderArgs& a) {
    REQUIRE(g && dec && opts, "null grid, decoder or options");
    REQUIRE(g->voxels && g->mask, "null grid arrays");
    REQUIRE(g->n > 0 && g->voxel_size > 0, "bad grid size");
    REQUIRE(g->channels == dec->desc.channels, "grid has %d channels, decoder expects %d",
            g->channels, dec->desc.channels);
    const int precision = opts->precision;
    REQUIRE(precision >= DIVER_FP32 && precision <= DIVER_FP16, "bad precision %d", precision);
    REQUIRE(opts->traversal == DIVER_WALK_OFFLINE || opts->traversal == DIVER_WALK_REALTIME,
            "bad traversal %d", opts->traversal);
    if (opts->traversal == DIVER_WALK_REALTIME) {
        REQUIRE(g->octrees, "the real-time walk needs the occupancy pyramid");
        REQUIRE(g->n >= 64 && g->n % 64 == 0 && g->n <= 4096,
                "the real-time walk needs n to be a multiple of 64 in [64, 4096]");
    }
    if (dec->variant == 5)
        return fail(DIVER_EUNSUPPORTED,
                    "the real-time blob decoder (pre-multiplied features) is for mlp_eval only");
    if (precision != DIVER_FP32 && dec->desc.n_pre != 0)
        return fail(DIVER_EUNSUPPORTED,
                    "the matrix-core decoder path supports decoders without pre-layers only");
    REQUIRE(precision == DIVER_FP32 || dec->packs_valid,
            "the decoder was updated from device memory; its BF16/FP16 operands are stale "
            "(re-create it from host parameters)");
    a.voxels = g->voxels;
    a.mask = g->mask;
    a.octrees = g->octrees;
    a.blocks = g->blocks;
    a.bricks = reinterpret_cast<const unsigned long long*>(g->bricks);
    if (g->blocks || g->bricks)
        REQUIRE(g->n % 8 == 0, "block and brick bits need n to be a multiple of 8");
    if (g->tail) REQUIRE(g->n % 8 == 0, "tail bits need n to be a multiple of 8");
    a.tail = g->tail;
    a.traversal = opts->traversal;
    a.n = g->n;
    a.channels = g->channels;
    a.xyzmin = g->xyzmin;
    a.voxel_size = g->voxel_size;
    a.W = dec->W;
    a.pack = precision == DIVER_BF16   ? dec->pack_bf16
             : precision == DIVER_FP16 ? dec->pack_f16
                                       : reinterpret_cast<const uint32_t*>(dec->pack_f32);
    a.pack16 = precision == DIVER_BF16   ? dec->pack16_bf16
               : precision == DIVER_FP16 ? dec->pack16_f16
                                         : nullptr;
    a.t_stop = opts->t_stop;
    a.alpha_skip = opts->alpha_skip;
    return DIVER_OK;
}

int diver_render_rays(const diver_grid* grid, const diver_decoder* dec,
                      const diver_render_opts* opts, int64_t n_rays, const float* o,
                      const float* d, float* rgba, int32_t* count, diver_stream_t stream) {
    diver::RenderArgs a{};
    int st = render_common(grid, dec, opts, a);
    if (st != DIVER_OK) return st;
    REQUIRE(n_rays >= 0, "n_rays < 0");
    REQUIRE(n_rays == 0 || (o && d && rgba), "null pointer");
    a.n_rays = n_rays;
    a.o = o;
    a.d = d;
    a.cam.width = 0;
    a.rgba = rgba;
    a.count = count;
    return dispatch_render(dec, a, opts->precision, S(stream), "render_rays");
}

int diver_render_rays_sparsity(const diver_grid* grid, const diver_decoder* dec,
                               const diver_render_opts* opts, int64_t n_rays, const float* o,
                               const float* d, float* rgba, int32_t* count, float* sparsity,
                               diver_stream_t stream) {
    diver::RenderArgs a{};
    int st = render_common(grid, dec, opts, a);
    if (st != DIVER_OK) return st;
    REQUIRE(n_rays >= 0, "n_rays < 0");
    REQUIRE(n_rays == 0 || (o && d && rgba && sparsity), "null pointer");
    a.n_rays = n_rays;
    a.o = o;
    a.d = d;
    a.cam.width = 0;
    a.rgba = rgba;
    a.count = count;
    if (!tiles_ok(dec, a))
        return fail(DIVER_EUNSUPPORTED,
                    "the sparsity output needs the tile kernel (block and brick bits, a decoder "
                    "without pre-layers)");
    diver::BatchArgs p{};
    p.g = a;
    p.n_req = 1;
    p.req[0].cam = a.cam;
    p.req[0].n_rays = n_rays;
    p.req[0].o = o;
    p.req[0].d = d;
    p.req[0].rgba = rgba;
    p.req[0].count = count;
    p.req[0].sparsity = sparsity;
    p.total = n_rays;
    p.sparsity = 1;
    return hip_status(diver::launch_render_tiles(dec->variant, p, opts->precision,
                                                 dec->density_ch0, S(stream)),
                      "render_rays_sparsity");
}

int diver_render_rays_segments(const diver_grid* grid, const diver_decoder* dec,
                               const diver_render_opts* opts, int64_t n_rays, const float* o,
                               const float* d, float* rgba, int32_t* count, float* sparsity,
                               float* segments, int seg_cap, diver_stream_t stream) {
    diver::RenderArgs a{};
    int st = render_common(grid, dec, opts, a);
    if (st != DIVER_OK) return st;
    REQUIRE(n_rays >= 0, "n_rays < 0");
    REQUIRE(n_rays == 0 || (o && d && rgba && count && segments), "null pointer");
    REQUIRE(seg_cap >= 1, "seg_cap must be positive");
    REQUIRE(opts->precision == DIVER_FP32 && opts->traversal == DIVER_WALK_OFFLINE &&
                opts->alpha_skip == 0.0f && opts->t_stop <= 0.0f,
            "segments are recorded by the FP32 offline walk with alpha_skip == 0 and no early "
            "termination (t_stop <= 0: the training forward)");
    a.n_rays = n_rays;
    a.o = o;
    a.d = d;
    a.cam.width = 0;
    a.rgba = rgba;
    a.count = count;
    if (!tiles_ok(dec, a) || !dec->density_ch0)
        return fail(DIVER_EUNSUPPORTED,
                    "segments need the tile kernel (block and brick bits, a decoder without "
                    "pre-layers) and the channel-0 density (w_s = e_0, b_s = 0)");
    diver::BatchArgs p{};
    p.g = a;
    p.n_req = 1;
    p.req[0].cam = a.cam;
    p.req[0].n_rays = n_rays;
    p.req[0].o = o;
    p.req[0].d = d;
    p.req[0].rgba = rgba;
    p.req[0].count = count;
    p.req[0].sparsity = sparsity;
    p.total = n_rays;
    p.sparsity = sparsity ? 1 : 0;
    p.segments = segments;
    p.seg_cap = seg_cap;
    return hip_status(diver::launch_render_tiles(dec->variant, p, opts->precision,
                                                 dec->density_ch0, S(stream)),
                      "render_rays_segments");
}

int diver_alpha_map(const diver_grid* grid, const diver_decoder* dec, int64_t n_rays,
                    const float* o, const float* d, int pad, float* alpha_map,
                    diver_stream_t stream) {
    diver::RenderArgs a{};
    const diver_render_opts opts{DIVER_FP32, DIVER_WALK_OFFLINE, 0.0f, 0.0f};
    int st = render_common(grid, dec, &opts, a);
    if (st != DIVER_OK) return st;
    REQUIRE(n_rays >= 0, "n_rays < 0");
    REQUIRE(n_rays == 0 || (o && d && alpha_map), "null pointer");
    REQUIRE(pad >= 0, "pad < 0");
    a.n_rays = n_rays;
    a.o = o;
    a.d = d;
    return hip_status(diver::launch_alpha_map(dec->variant, a, pad, alpha_map, dec->pack_pre,
                                              S(stream)),
                      "alpha_map");
}

// The Blender camera (utils/dataset/ray_utils.py): fx = fy = focal, principal point at the centre.
static void blender_camera(diver::CameraDesc& cam, const float* c2w, int width, int height,
                           float focal) {
    std::memcpy(cam.c2w, c2w, sizeof(cam.c2w));
    cam.width = width;
    cam.height = height;
    cam.fx = cam.fy = focal;
    cam.cx = 0.5f * (float)width;
    cam.cy = 0.5f * (float)height;
}

int diver_render_camera(const diver_grid* grid, const diver_decoder* dec,
                        const diver_render_opts* opts, const float* c2w_host, int width,
                        int height, float focal, float* rgba, int32_t* count,
                        diver_stream_t stream) {
    diver::RenderArgs a{};
    int st = render_common(grid, dec, opts, a);
    if (st != DIVER_OK) return st;
    REQUIRE(c2w_host && rgba, "null pointer");
    REQUIRE(width > 0 && height > 0 && focal > 0, "bad camera");
    blender_camera(a.cam, c2w_host, width, height, focal);
    int64_t tiles = (int64_t)((width + 7) / 8) * ((height + 7) / 8);
    a.n_rays = tiles * 64;
    a.rgba = rgba;
    a.count = count;
    return dispatch_render(dec, a, opts->precision, S(stream), "render_camera");
}

int diver_render_batch(const diver_grid* grid, const diver_decoder* dec,
                       const diver_render_opts* opts, const diver_request* reqs, int n_req,
                       diver_stream_t stream) {
    diver::RenderArgs a{};
    int st = render_common(grid, dec, opts, a);
    if (st != DIVER_OK) return st;
    REQUIRE(n_req >= 0 && n_req <= DIVER_MAX_REQUESTS, "n_req must be in [0, %d]",
            DIVER_MAX_REQUESTS);
    REQUIRE(n_req == 0 || reqs, "null request table");
    if (!tiles_ok(dec, a))
        return fail(DIVER_EUNSUPPORTED,
                    "batched rendering needs block and brick bits (n <= 512), a decoder without "
                    "pre-layers and (n+1)^3 * channels < 2^32");
    diver::BatchArgs p{};
    p.g = a;
    int64_t total = 0;
    for (int i = 0; i < n_req; ++i) {
        const diver_request& r = reqs[i];
        diver::ReqDesc& q = p.req[i];
        REQUIRE(r.rgba || (r.width == 0 && r.n_rays == 0), "request %d: null rgba", i);
        if (r.width > 0) {
            REQUIRE(r.height > 0 && r.focal > 0 && r.fy >= 0, "request %d: bad camera", i);
            blender_camera(q.cam, r.c2w, r.width, r.height, r.focal);
            if (r.fy > 0) {  // a general pinhole camera (Tanks and Temples, BlendedMVS)
                q.cam.fy = r.fy;
                q.cam.cx = r.cx;
                q.cam.cy = r.cy;
            }
            q.n_rays = (int64_t)((r.width + 7) / 8) * ((r.height + 7) / 8) * 64;
        } else {
            REQUIRE(r.width == 0, "request %d: negative width", i);
            REQUIRE(r.n_rays >= 0, "request %d: n_rays < 0", i);
            REQUIRE(r.n_rays == 0 || (r.o && r.d), "request %d: null rays", i);
            q.cam.width = 0;
            q.n_rays = r.n_rays;
            q.o = r.o;
            q.d = r.d;
        }
        q.rgba = r.rgba;
        q.count = r.count;
        q.ray_start = total;
        total += q.n_rays;
    }
    p.n_req = n_req;
    p.total = total;
    return hip_status(diver::launch_render_tiles(dec->variant, p, opts->precision,
                                                 dec->density_ch0, S(stream)),
                      "render_batch");
}

int diver_render_rows(const diver_grid* grid, const diver_decoder* dec,
                      const diver_render_opts* opts, int64_t n_rays, const float* o,
                      const float* d, const int32_t* count, const int64_t* row_offset,
                      float* rows, diver_stream_t stream) {
    diver::RenderArgs a{};
    int st = render_common(grid, dec, opts, a);
    if (st != DIVER_OK) return st;
    if (dec->desc.n_pre != 0)
        return fail(DIVER_EUNSUPPORTED, "rows are written for decoders without pre-layers");
    REQUIRE(opts->alpha_skip == 0.0f, "the adjoint needs alpha_skip == 0");
    REQUIRE(n_rays >= 0, "n_rays < 0");
    REQUIRE(n_rays == 0 || (o && d && count && row_offset && rows), "null pointer");
    a.n_rays = n_rays;
    a.o = o;
    a.d = d;
    return hip_status(diver::launch_rows(dec->variant, a, dec->density_ch0, count, row_offset, rows, S(stream)),
                      "render_rows");
}

int diver_render_rows_segments(const diver_grid* grid, const diver_decoder* dec,
                               const diver_render_opts* opts, int64_t n_rays, const int32_t* count,
                               const int64_t* row_offset, const float* segments, int seg_cap,
                               float* rows, int64_t* ray, diver_stream_t stream) {
    diver::RenderArgs a{};
    int st = render_common(grid, dec, opts, a);
    if (st != DIVER_OK) return st;
    if (dec->desc.n_pre != 0)
        return fail(DIVER_EUNSUPPORTED, "rows are written for decoders without pre-layers");
    REQUIRE(opts->alpha_skip == 0.0f, "the adjoint needs alpha_skip == 0");
    REQUIRE(n_rays >= 0, "n_rays < 0");
    REQUIRE(seg_cap >= 1, "seg_cap must be positive");
    REQUIRE(n_rays == 0 || (count && row_offset && segments && rows), "null pointer");
    a.n_rays = n_rays;
    return hip_status(diver::launch_rows_from_segments(dec->variant, a, dec->density_ch0, count,
                                                       row_offset, segments, seg_cap, rows, ray,
                                                       S(stream)),
                      "render_rows_segments");
}

int diver_adjoint_decoder_forward(const diver_decoder* dec, int64_t n_rows, const float* rows,
                                  const int64_t* ray, const float* d, float* r1, float* r2,
                                  float* color, float* weight, diver_stream_t stream) {
    REQUIRE(dec, "null decoder");
    REQUIRE(dec->pack_adj && dec->pack_f32,
            "the decoder adjoint needs a decoder without pre-layers, one post layer and channels "
            "and hidden width that are multiples of 32");
    REQUIRE(n_rows >= 0, "n_rows < 0");
    if (n_rows == 0) return DIVER_OK;
    REQUIRE(rows && ray && d && r1 && r2 && color && weight, "null pointer");
    diver::AdjArgs p{};
    p.n_rows = n_rows;
    p.n_valid = n_rows;
    p.rows = rows;
    p.row_stride = 9 + dec->desc.channels;
    p.ray = ray;
    p.d = d;
    p.W = dec->W;
    p.pack = dec->pack_f32;
    p.pack_adj = dec->pack_adj;
    p.r1 = r1;
    p.r2 = r2;
    p.color = color;
    p.weight = weight;
    return hip_status(diver::launch_adjoint(dec->variant, false, p, S(stream)),
                      "adjoint_decoder_forward");
}

int diver_adjoint_decoder_backward(const diver_decoder* dec, int64_t n_rows, int64_t n_valid,
                                   const float* rows, const int64_t* ray, const float* grad_rgb,
                                   const float* suffix, const float* sparsity_scale,
                                   const float* r1, const float* r2, const float* color,
                                   const float* weight, float* d_feature, float* d_out, float* d_s,
                                   float* d_r1, float* d_r2, diver_stream_t stream) {
    REQUIRE(dec, "null decoder");
    REQUIRE(dec->pack_adj && dec->pack_f32,
            "the decoder adjoint needs a decoder without pre-layers, one post layer and channels "
            "and hidden width that are multiples of 32");
    REQUIRE(n_rows >= 0 && n_valid >= 0 && n_valid <= n_rows, "need 0 <= n_valid <= n_rows");
    if (n_rows == 0) return DIVER_OK;
    REQUIRE(rows && ray && grad_rgb && r1 && r2 && color && weight && d_out && d_s && d_r1 && d_r2,
            "null pointer");
    REQUIRE(n_valid == 0 || (suffix && d_feature), "null pointer");
    diver::AdjArgs p{};
    p.n_rows = n_rows;
    p.n_valid = n_valid;
    p.rows = rows;
    p.row_stride = 9 + dec->desc.channels;
    p.ray = ray;
    p.W = dec->W;
    p.pack = dec->pack_f32;
    p.pack_adj = dec->pack_adj;
    p.r1 = const_cast<float*>(r1);
    p.r2 = const_cast<float*>(r2);
    p.color = const_cast<float*>(color);
    p.weight = const_cast<float*>(weight);
    p.g = grad_rgb;
    p.suffix = suffix;
    p.sparsity_scale = sparsity_scale;
    p.dF = d_feature;
    p.dout = d_out;
    p.ds = d_s;
    p.dr1 = d_r1;
    p.dr2 = d_r2;
    return hip_status(diver::launch_adjoint(dec->variant, true, p, S(stream)),
                      "adjoint_decoder_backward");
}

int64_t diver_adjoint_weight_grads_workspace(const diver_decoder* dec, int64_t n_rows) {
    if (!dec || n_rows < 0 || dec->variant < 0 || dec->variant > 2 || dec->desc.n_pre != 0)
        return -1;
    const int tiles = diver::wgrad_tiles(dec->desc.channels, dec->desc.hidden);
    return (int64_t)sizeof(float) * diver::wgrad_groups(n_rows) * tiles * 32;
}

int diver_adjoint_weight_grads(const diver_decoder* dec, int64_t n_rows, const float* rows,
                               const int64_t* ray, const float* d, const float* r1, const float* r2,
                               const float* d_out, const float* d_s, const float* d_r1,
                               const float* d_r2, float* grad, float* workspace,
                               int64_t workspace_bytes, diver_stream_t stream) {
    REQUIRE(dec, "null decoder");
    REQUIRE(dec->variant >= 0 && dec->variant <= 2 && dec->desc.n_pre == 0,
            "the weight gradients need a decoder without pre-layers, one post layer and "
            "channels/hidden 32/32, 32/64 or 64/64");
    REQUIRE(n_rows >= 0, "n_rows < 0");
    REQUIRE(grad, "null gradient pointer");
    const int64_t need = diver_adjoint_weight_grads_workspace(dec, n_rows);
    REQUIRE(workspace && workspace_bytes >= need,
            "workspace of %lld bytes needed (diver_adjoint_weight_grads_workspace), %lld given",
            (long long)need, (long long)workspace_bytes);
    REQUIRE(n_rows == 0 || (rows && ray && d && r1 && r2 && d_out && d_s && d_r1 && d_r2),
            "null pointer");
    REQUIRE(((reinterpret_cast<uintptr_t>(r1) | reinterpret_cast<uintptr_t>(r2) |
              reinterpret_cast<uintptr_t>(d_r1) | reinterpret_cast<uintptr_t>(d_r2)) & 15) == 0,
            "r1, r2, d_r1 and d_r2 must be 16-byte aligned");
    diver::WgradArgs p{};
    p.n_rows = n_rows;
    p.rows = rows;
    p.ray = ray;
    p.d = d;
    p.r1 = r1;
    p.r2 = r2;
    p.dout = d_out;
    p.ds = d_s;
    p.dr1 = d_r1;
    p.dr2 = d_r2;
    p.partial = workspace;
    return hip_status(diver::launch_weight_grads(dec->variant, p, grad, S(stream)),
                      "adjoint_weight_grads");
}

int diver_adjoint_weight_grads_noalloc(const diver_decoder* dec, int64_t n_rows, const float* rows,
                                       const int64_t* ray, const float* d, const float* r1,
                                       const float* r2, const float* d_out, const float* d_s,
                                       const float* d_r1, const float* d_r2, float* grad,
                                       diver_stream_t stream) {
    REQUIRE(dec, "null decoder");
    REQUIRE(n_rows >= 0, "n_rows < 0");
    const int64_t need = diver_adjoint_weight_grads_workspace(dec, n_rows);
    REQUIRE(need >= 0,
            "the weight gradients need a decoder without pre-layers, one post layer and "
            "channels/hidden 32/32, 32/64 or 64/64");
    float* ws = nullptr;
    int64_t cap = 0;
    const hipError_t e = diver::wgrad_scratch(&ws, &cap);
    if (e != hipSuccess) return hip_status(e, "adjoint_weight_grads_noalloc: scratch address");
    REQUIRE(need <= cap, "the partial sums need %lld bytes, above the %lld-byte scratch array",
            (long long)need, (long long)cap);
    return diver_adjoint_weight_grads(dec, n_rows, rows, ray, d, r1, r2, d_out, d_s, d_r1, d_r2,
                                      grad, ws, cap, stream);
}

int diver_positional_encoding(int64_t n, const float* d, float* pe, diver_stream_t stream) {
    REQUIRE(n >= 0, "n < 0");
    if (n == 0) return DIVER_OK;
    REQUIRE(d && pe, "null pointer");
    return hip_status(diver::launch_positional_encoding(n, d, pe, S(stream)), "positional_encoding");
}

int diver_ray_suffix(int64_t n_rays, const int64_t* row_offset, const int32_t* count,
                     const float* weight, const float* color, const float* background,
                     float* suffix, diver_stream_t stream) {
    REQUIRE(n_rays >= 0, "n_rays < 0");
    if (n_rays == 0) return DIVER_OK;
    REQUIRE(row_offset && count && weight && color && suffix, "null pointer");
    return hip_status(diver::launch_ray_suffix(n_rays, row_offset, count, weight, color, background,
                                               suffix, S(stream)),
                      "ray_suffix");
}

int diver_camera_rays(const float* c2w_host, int width, int height, float focal, float* o,
                      float* d, diver_stream_t stream) {
    REQUIRE(c2w_host && o && d, "null pointer");
    REQUIRE(width > 0 && height > 0 && focal > 0, "bad camera");
    diver::CameraDesc cam{};
    blender_camera(cam, c2w_host, width, height, focal);
    return hip_status(diver::launch_camera_rays(cam, o, d, S(stream)), "camera_rays");
}

int diver_camera_rays_pinhole(const float* c2w_host, int width, int height, float fx, float fy,
                              float cx, float cy, float* o, float* d, diver_stream_t stream) {
    REQUIRE(c2w_host && o && d, "null pointer");
    REQUIRE(width > 0 && height > 0 && fx > 0 && fy > 0, "bad camera");
    diver::CameraDesc cam{};
    blender_camera(cam, c2w_host, width, height, fx);
    cam.fy = fy;
    cam.cx = cx;
    cam.cy = cy;
    return hip_status(diver::launch_camera_rays(cam, o, d, S(stream)), "camera_rays_pinhole");
}

}  // extern "C"
