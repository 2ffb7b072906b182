// Internal launcher interface between the C-ABI (capi.cpp) and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace diver {

struct CameraDesc {
    float c2w[12];  // row-major 3x4 camera-to-world
    int width;      // 0 = explicit rays
    int height;
    float fx, fy;   // focal lengths in pixels
    float cx, cy;   // principal point in pixels (width / 2, height / 2 for the Blender camera)
};

struct RenderArgs {
    const float* voxels;   // [(N+1)^3][C] fp32
    const uint8_t* mask;   // [N^3]
    const uint8_t* octrees;  // occupancy pyramid (block sizes 64, 16, 4), real-time traversal
    const uint32_t* blocks;  // 8^3-block occupancy bits (optional)
    const unsigned long long* bricks;  // bit-packed mask, 64 bits per 4^3 brick (optional)
    const uint32_t* tail;  // tail bits: 16^3 cells with an occupied voxel around them (optional)
    int traversal;         // 0: masked_intersect walk (exit > N); 1: real-time walk
    int n;
    int channels;
    float xyzmin;
    float voxel_size;
    const float* W;        // canonical decoder params (fp32, device)
    const uint32_t* pack;  // matrix-core A operands for the chosen precision (device)
    const uint32_t* pack16;  // BF16/FP16 A operands in the 16x16x32 layout (tile kernel)
    int64_t n_rays;        // explicit: number of rays; camera: tiles*64
    const float* o;
    const float* d;
    CameraDesc cam;
    float t_stop;
    float alpha_skip;
    float* rgba;           // [rays][4]: accumulated rgb (no background) and transmittance
    int32_t* count;        // [rays] intervals consumed (optional)
};

// One request of a batched render: a camera view (width > 0) or explicit rays.  Its rays are
// [ray_start, ray_start + n_rays) of the launch's ray space (camera: tiles * 64 in tile order).
struct ReqDesc {
    CameraDesc cam;
    int64_t ray_start;
    int64_t n_rays;
    const float* o;
    const float* d;
    float* rgba;
    int32_t* count;
    float* sparsity;  // [rays] sum of log(1 + 2 s^2) over the evaluated intervals (optional)
};

constexpr int kMaxRequests = 16;

// A launch over a table of requests (diver_render_batch; a single render is a table of one).
struct BatchArgs {
    RenderArgs g;                  // grid, decoder, options (ray fields unused)
    ReqDesc req[kMaxRequests];
    int n_req;
    int64_t total;                 // ray ids over all requests
    int rays_per_wave;             // ids per wavefront (64, or fewer for small launches; set by
                                   // the launcher)
    int sparsity;                  // some request has a sparsity output
    float* segments;               // [rays][seg_cap][6] entry and exit of each evaluated interval
                                   // (single-request FP32 offline launches; null = not recorded)
    int seg_cap;
};

// Rays per wavefront for a launch of `total` rays: 64 when the launch fills the GPU, else halved
// (down to 4) until it makes about 2048 waves.  A wave's time is set by its longest ray and by the
// shading work of its own rays, so a small launch finishes sooner as more, lighter waves spread
// over the 256 CUs (8192 training rays: 4 rays per wave, 0.64 -> 0.57 ms for the forward; 2 per
// wave is slower again).  A ray's result does not depend on the rays it shares a wave with.
inline int rays_per_wave(int64_t total) {
    constexpr int64_t kTargetWaves = 2048;
    const int64_t want = total / kTargetWaves;
    int rpw = 64;
    while (rpw > 4 && rpw > want) rpw >>= 1;
    return rpw;
}

hipError_t launch_ray_voxel_intersect(int64_t batch, int max_n, float xyzmin, float voxel_num,
                                      float voxel_size, const float* o, const float* v,
                                      float* isect, int32_t* num, float* tns, hipStream_t s);
hipError_t launch_masked_intersect(int64_t batch, int max_n, float xyzmin, float voxel_num,
                                   float voxel_size, float mask_scale, const float* o,
                                   const float* v, const uint8_t* mask, float* isect, int32_t* num,
                                   float* tns, hipStream_t s);
hipError_t launch_aabb_intersect(int64_t batch, float* coords, const float* v, const float* center,
                                 uint8_t* finish, float* rgba, const uint8_t* octrees,
                                 int voxel_num, hipStream_t s);
hipError_t launch_ray_march(int64_t batch, float* coords, const float* v, const uint8_t* mask,
                            const uint8_t* finish, int voxel_num, int64_t stride, hipStream_t s);
hipError_t launch_integrate(const float* voxels, int n1, int C, int64_t K, const float* ci,
                            const float* co, float* out, hipStream_t s);
hipError_t launch_integrate_backward(int n1, int C, int64_t K, const float* ci, const float* co,
                                     const float* gout, float* gvox, hipStream_t s);
hipError_t launch_composite(int64_t batch, int m, const float* sigma, const float* color,
                            const uint8_t* mask, int white_back, float* rgb, float* weight,
                            hipStream_t s);
hipError_t launch_composite_backward(int64_t batch, int m, const float* sigma, const float* color,
                                     const uint8_t* mask, int white_back, const float* grad_rgb,
                                     const float* grad_weight, float* grad_sigma,
                                     float* grad_color, hipStream_t s);
hipError_t launch_mlp_eval(int variant, int64_t batch, const float* coord, const float* feats,
                           const int32_t* vmap, int map_size, const float* v, float* rgba,
                           uint8_t* finish, const float* W, float alpha_skip, float t_stop,
                           hipStream_t s);
hipError_t launch_render(int variant, const RenderArgs& a, int prec, hipStream_t s);
hipError_t launch_alpha_map(int variant, const RenderArgs& a, int pad, float* map,
                            const float* pack_pre, hipStream_t s);
// Decoder and compositing adjoint over the recompute pass's rows (adjoint_mlp.hip).
struct AdjArgs {
    int64_t n_rows;          // rows (any rows past n_valid are zero padding)
    int64_t n_valid;         // composited intervals
    const float* rows;       // [n_rows][row_stride]: p_in, p_out, s, T_before, alpha, F
    int row_stride;          // 9 + C
    const int64_t* ray;      // [n_rows] ray of each row
    const float* d;          // [R][3] ray directions (forward: the view term)
    const float* W;          // canonical decoder parameters
    const float* pack;       // FP32 forward operands (pack_f32)
    const float* pack_adj;   // FP32 transposed operands (pack_adj)
    float* r1;               // [n_rows][H] relu(y1)
    float* r2;               // [n_rows][H] relu(y2)
    float* color;            // [n_rows][3]
    float* weight;           // [n_rows] T_before * alpha
    const float* g;          // [R][3] dL/d(rgb)
    const float* suffix;     // [n_valid][3] per-ray suffix sums of w c (+ background)
    const float* sparsity_scale;  // device scalar g_reg / (B M), or null
    float* dF;               // [n_valid][C]
    float* dout;             // [n_rows][3]
    float* ds;               // [n_rows]
    float* dr1;              // [n_rows][H]
    float* dr2;              // [n_rows][H]
};
int64_t pack_adj_size(int C, int H, int npost);
hipError_t launch_pack_adj(int C, int H, int npost, const float* W, float* out, hipStream_t s);
hipError_t launch_adjoint(int variant, bool backward, const AdjArgs& p, hipStream_t s);

// Weight gradients of the decoder adjoint over n_rows rows (variants 0-2: C/H 32/32, 32/64, 64/64).
struct WgradArgs {
    int64_t n_rows;
    int64_t rows_per_group;  // set by the launcher
    const float* rows;       // [n_rows][9 + C]: F at 9
    const int64_t* ray;      // [n_rows]
    const float* d;          // [R][3]
    const float* r1;         // [n_rows][H]
    const float* r2;         // [n_rows][H]
    const float* dout;       // [n_rows][3]
    const float* ds;         // [n_rows]
    const float* dr1;        // [n_rows][H]
    const float* dr2;        // [n_rows][H]
    float* partial;          // [groups][tiles][32] per-workgroup partial sums
};
constexpr int kWgradMaxGroups = 512;
int wgrad_tiles(int C, int H);                 // tiles per partial (0: unsupported shape)
int64_t wgrad_groups(int64_t n_rows);          // workgroups of the first pass
hipError_t launch_weight_grads(int variant, const WgradArgs& p, float* grad, hipStream_t s);
// Device address and size of the file-scope partial array of the allocation-free weight
// gradients on the current device (looked up on the first call per device).
hipError_t wgrad_scratch(float** ptr, int64_t* bytes);
hipError_t launch_positional_encoding(int64_t R, const float* d, float* pe, hipStream_t s);

// FP32 matrix-core A operands of a decoder's pre-layers (the occupancy kernel's density MLP),
// packed on the device from its canonical parameters W: pack_pre_f32_size floats.
int64_t pack_pre_f32_size(int C, int H, int npre);
hipError_t launch_pack_pre_f32(int C, int H, int npre, const float* W, float* out, hipStream_t s);
hipError_t launch_corner_keys(int n1, int64_t K, const float* ci, const float* co, int64_t* keys,
                              hipStream_t s);
hipError_t launch_integrate_table(const float* table, int C, int64_t K, const int64_t* crow,
                                  const float* ci, const float* co, float* out, hipStream_t s);
hipError_t launch_integrate_table_backward(int C, int64_t K, const int64_t* crow, const float* ci,
                                           const float* co, const float* gout, float* gtable,
                                           hipStream_t s);
hipError_t launch_render_tiles(int variant, const BatchArgs& p, int prec, bool dch0,
                               hipStream_t s);
bool tile_fits(int n);  // the grid's block bitmap fits the tile kernel's LDS
// FP32 matrix-core A operands of a decoder without pre-layers (C channels, hidden H, npost post
// layers), packed on the device from its canonical parameters W: pack_f32_size floats.
int64_t pack_f32_size(int C, int H, int npost);
hipError_t launch_pack_f32(int C, int H, int npost, const float* W, float* out, hipStream_t s);
hipError_t launch_block_bits(const uint8_t* mask, int n, uint32_t* bits, hipStream_t s);
hipError_t launch_adam(float* p, float* g, float* m, float* v, int64_t n, double b1, double b2,
                       float eps, float wd, float step_size, float bc2_sqrt, bool clear,
                       hipStream_t st);
hipError_t launch_adam_rows(float* p, float* g, float* m, float* v, int64_t n, const int64_t* rows,
                            int64_t n_rows, int row_len, double b1, double b2, float eps,
                            float step_size, float bc2_sqrt, bool clear, hipStream_t st);
hipError_t launch_brick_bits(const uint8_t* mask, int n, unsigned long long* bricks,
                             hipStream_t s);
hipError_t launch_walk(const RenderArgs& a, int max_n, float* isect, int32_t* num, hipStream_t s);
hipError_t launch_tail_bits(const uint32_t* blocks, int n, uint32_t* tail, hipStream_t s);
hipError_t launch_ray_suffix(int64_t R, const int64_t* off, const int32_t* cnt, const float* w,
                             const float* c, const float* bg, float* out, hipStream_t s);
// The rows from the forward's recorded segments instead of a re-walk (the same rows, bit for bit).
hipError_t launch_rows_from_segments(int variant, const RenderArgs& a, bool dch0, const int32_t* count,
                                     const int64_t* off, const float* segments, int seg_cap,
                                     float* rows, int64_t* ray, hipStream_t s);
hipError_t launch_rows(int variant, const RenderArgs& a, bool dch0, const int32_t* count,
                       const int64_t* off, float* rows, hipStream_t s);
hipError_t launch_camera_rays(const CameraDesc& cam, float* o, float* d, hipStream_t s);

}  // namespace diver
