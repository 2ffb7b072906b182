/*
 * libdiver: MI355X (gfx950) implementation of DIVeR's render-and-train hot path behind a C ABI.
 *
 * Conventions for every entry point:
 *   - pointers are device pointers unless the parameter name ends in _host;
 *   - work is enqueued on `stream` (a hipStream_t; NULL = the null stream) and the call returns
 *     without a device-wide synchronisation;
 *   - the return value is a diver_status (DIVER_OK = 0); diver_last_error() describes the last
 *     failure of the calling thread; no entry point aborts the process;
 *   - every output element is written (padding included), so outputs need no pre-fill.
 * Coordinates are grid units ([0, N] per axis), features are [z][y][x][c] with (N+1)^3 corners,
 * masks are [z][y][x] bytes (0/1).
 */
#ifndef DIVER_H
#define DIVER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* diver_stream_t;

enum diver_status {
    DIVER_OK = 0,
    DIVER_EINVAL = 1,       /* bad argument (null pointer, size, unsupported shape)          */
    DIVER_ENOMEM = 2,       /* device allocation failed                                       */
    DIVER_EHIP = 3,         /* HIP runtime error (launch or copy)                             */
    DIVER_EUNSUPPORTED = 4, /* combination not compiled into the library                      */
    DIVER_EBUDGET = 5,      /* a render server's memory budget can never hold the scene/frame    */
    DIVER_ETIMEDOUT = 6,    /* a wait ended before the frame was rendered                        */
    DIVER_ESTATE = 7        /* the server has failed or is closing; the frame was not rendered    */
};

const char* diver_version(void);
const char* diver_last_error(void);

/* ---- traversal (training / offline) ------------------------------------------------------- */

/* Every voxel-boundary crossing of each ray.  Replaces the pybind11 op
 * ray_voxel_intersect(o, v, xyzmin, xyzmax, voxel_num, voxel_size)
 * (utils/ray_voxel_intersection/include/ray_voxel_intersect.h:5-9, kernel
 * src/ray_voxel_intersect.cu:29-151).  isect [batch][max_n][3], tns [batch][max_n]; entries past
 * the count are 0 / -1.  isect and tns may be NULL to compute counts only.  num[b] is the true
 * count even when it exceeds max_n (entries past max_n are dropped). */
int diver_ray_voxel_intersect(int64_t batch, int max_n, float xyzmin, float xyzmax,
                              float voxel_num, float voxel_size, const float* o, const float* v,
                              float* isect, int32_t* num, float* tns, diver_stream_t stream);

/* (entry, exit) of every occupied voxel along each ray.  Replaces masked_intersect(o, v, mask,
 * xyzmin, xyzmax, voxel_num, voxel_size, mask_scale)
 * (utils/ray_voxel_intersection/include/ray_voxel_intersect.h:11-15, kernel
 * src/ray_voxel_intersect.cu:192-319).  mask [M^3] with M = int(voxel_num * mask_scale).
 * isect [batch][max_n][6], tns [batch][max_n] = exit distance; NULL isect/tns = counts only. */
int diver_masked_intersect(int64_t batch, int max_n, float xyzmin, float xyzmax, float voxel_num,
                           float voxel_size, float mask_scale, const float* o, const float* v,
                           const uint8_t* mask, float* isect, int32_t* num, float* tns,
                           diver_stream_t stream);

/* ---- traversal (real time) ------------------------------------------------------------------ */

/* Box entry + coarse-to-fine octree descent to the first occupied N/4 block; initialises rgba to
 * (0,0,0,1) and finish.  Replaces aabb_intersect(o, v, center, finish, rgba, octrees, voxel_num)
 * (utils/ray_march/include/ray_march.h:5-10, kernel src/ray_march.cu:24-149).  coords
 * [8][batch][6] (slot 7's exit receives the start point, -1 on a miss); center [3] device, grid
 * units; octrees = max-pooled masks at block sizes 64, 16, 4 concatenated coarse first.
 * voxel_num must be a positive multiple of 64 (the reference compiles 128 and 256 only). */
int diver_aabb_intersect(int64_t batch, float* coords, const float* v, const float* center,
                         uint8_t* finish, float* rgba, const uint8_t* octrees, int voxel_num,
                         diver_stream_t stream);

/* One march round: up to 8 occupied (entry, exit) pairs per unfinished ray into coords slots
 * 0..7, -1 terminator.  Replaces ray_march(o, v, mask, finish, voxel_num)
 * (utils/ray_march/include/ray_march.h:12-16, kernel src/ray_march.cu:151-230).  thread_stride
 * is the reference's grid-stride width (262144 = 256 x 1024; <= 0 means batch), which decides
 * which rays a thread skips after meeting a negative resume point. */
int diver_ray_march(int64_t batch, float* coords, const float* v, const uint8_t* mask,
                    const uint8_t* finish, int voxel_num, int64_t thread_stride,
                    diver_stream_t stream);

/* ---- segment integral ----------------------------------------------------------------------- */

/* Mean of the trilinear feature field over each segment.  Replaces integrate(voxels, coord_i,
 * coord_o) (utils/integrator.py:62-113).  voxels [n1][n1][n1][channels], coords [k][3],
 * feature [k][channels]; channels must be a multiple of 4. */
int diver_integrate(const float* voxels, int n1, int channels, int64_t k, const float* coord_in,
                    const float* coord_out, float* feature, diver_stream_t stream);

/* Adjoint of diver_integrate: grad_voxels += d feature / d voxels ^T grad_feature (atomic
 * scatter-add; the reference gets this from autograd through integrator.py's gathers). */
int diver_integrate_backward(int n1, int channels, int64_t k, const float* coord_in,
                             const float* coord_out, const float* grad_feature,
                             float* grad_voxels, diver_stream_t stream);

/* Implicit grid (utils/integrator.py:3-60, integrate_mlp): the features of a voxel's corners come
 * from an MLP evaluated once per distinct corner.  diver_corner_keys writes, per row, the keys
 * z n1^2 + y n1 + x of the 8 corners of its integral (corner j = (j&1, (j>>1)&1, j>>2); upper
 * corner clamped to n1 - 1 as integrator.py:20), keys [k][8] int64.  After the caller has
 * evaluated the MLP at the distinct keys, diver_integrate_table integrates row r from
 * table [U][channels] rows corner_row[8 r + j] with the same weights as diver_integrate;
 * diver_integrate_table_backward adds d feature / d table ^T grad_feature into grad_table
 * (atomic). */
int diver_corner_keys(int n1, int64_t k, const float* coord_in, const float* coord_out,
                      int64_t* keys, diver_stream_t stream);
int diver_integrate_table(const float* table, int channels, int64_t k, const int64_t* corner_row,
                          const float* coord_in, const float* coord_out, float* feature,
                          diver_stream_t stream);
int diver_integrate_table_backward(int channels, int64_t k, const int64_t* corner_row,
                                   const float* coord_in, const float* coord_out,
                                   const float* grad_feature, float* grad_table,
                                   diver_stream_t stream);

/* ---- compositing ------------------------------------------------------------------------------ */

/* alpha = 1 - exp(-relu(sigma) * mask); weight = alpha * exclusive cumprod(1 - alpha);
 * rgb = sum weight * color (+ 1 - sum weight with white_back).  Replaces DIVeR.render
 * (model/diver.py:183-207).  sigma/mask [batch][m], color [batch][m][3], rgb [batch][3]. */
int diver_composite(int64_t batch, int m, const float* sigma, const float* color,
                    const uint8_t* mask, int white_back, float* rgb, float* weight,
                    diver_stream_t stream);

/* Adjoint of diver_composite (grad_weight may be NULL). */
int diver_composite_backward(int64_t batch, int m, const float* sigma, const float* color,
                             const uint8_t* mask, int white_back, const float* grad_rgb,
                             const float* grad_weight, float* grad_sigma, float* grad_color,
                             diver_stream_t stream);

/* ---- decoder ---------------------------------------------------------------------------------- */

/* Canonical decoder (the reference's mlp1/mlp2 after convert.py:70-77's layer fusion):
 *   z = F;  z = relu(W_i z + b_i) for n_pre layers;  sigma = act(w_s.z + b_s)
 *   y = W_c z + W_v PE(d) + b_c;  y = W_j relu(y) + b_j for n_post layers
 *   rgb = sigmoid(W_o relu(y) + b_o)
 * Parameter layout (float32, row-major [out][in], Z = n_pre ? hidden : channels, P = 27):
 *   pre 0: pre_identity ? b[H] : W[H][C], b[H];  pre i>0: W[H][H], b[H]
 *   w_s[Z], b_s, W_c[H][Z], W_v[H][P], b_c[H];  post: W[H][H], b[H];  W_o[3][H], b_o[3]
 * PE(d) = [d, sin(pi d), cos(pi d), ..., sin(8 pi d), cos(8 pi d)] (model/mlps.py:12-18). */
typedef struct diver_decoder_desc {
    int channels;
    int hidden;
    int n_pre;
    int pre_identity;
    int density_act; /* 0 = relu (model/diver.py:195), 1 = softplus */
    int n_post;
} diver_decoder_desc;

typedef struct diver_decoder diver_decoder;

int diver_decoder_create(const diver_decoder_desc* desc, const float* params_host,
                         int64_t n_params, diver_decoder** out);
int diver_decoder_destroy(diver_decoder* dec);
/* Replace the parameters from device memory (asynchronous copy on `stream`, e.g. after an
 * optimizer step).  The FP32 matrix-core operands are repacked on the device on the same stream;
 * only the FP32 decoder path may be used afterwards (the BF16/FP16 operands are derived on the
 * host at creation). */
int diver_decoder_update(diver_decoder* dec, const float* params, int64_t n_params,
                         diver_stream_t stream);
/* diver_decoder_update for parameters [first, first + count) only: params points at the count new
 * values.  A range that leaves the density row (w_s, b_s: the first channels + 1 parameters)
 * untouched keeps the channel-0 density shortcut of a decoder created with sigma = act(F[0]), so
 * training a colour MLP over a fixed density head keeps the cheaper density path. */
int diver_decoder_update_range(diver_decoder* dec, const float* params, int64_t first,
                               int64_t count, diver_stream_t stream);
int64_t diver_decoder_param_count(const diver_decoder_desc* desc);

/* Real-time decode + composite of one march round.  Replaces mlp_eval(rgba, coord, voxels, v,
 * mask) (utils/mlp_evaluation/include/mlp_eval.h:5-11, kernels
 * utils/mlp_evaluation{,64}/src/mlp_eval.cu:32-231).  The reference's upload_weight
 * (mlp_eval.h:13-17) installs the decoder and the corner map as process-global state; here they
 * are explicit: `dec` plus `voxel_map` (device int32 [(map_size)^3], rows of `features`).
 * features [rows][channels]; alpha_skip / t_stop are the reference's 1e-2 / 1e-2. */
int diver_mlp_eval(const diver_decoder* dec, int64_t batch, const float* coord,
                   const float* features, const int32_t* voxel_map, int map_size, const float* v,
                   float* rgba, uint8_t* finish, float alpha_skip, float t_stop,
                   diver_stream_t stream);

/* ---- fused renderer --------------------------------------------------------------------------- */

typedef struct diver_grid {
    const float* voxels;    /* [(n+1)^3][channels] */
    const uint8_t* mask;    /* [n^3] */
    const uint8_t* octrees; /* occupancy pyramid for the real-time walk (block sizes 64, 16, 4,
                               coarse first, eval_rt.py:236-239); NULL for the offline walk */
    const uint32_t* blocks; /* 8^3-block occupancy bits from diver_block_bits, or NULL; lets
                               empty blocks skip the fine mask (no effect on results) */
    const uint64_t* bricks; /* bit-packed mask, one 64-bit word per 4^3 brick, from
                               diver_brick_bits, or NULL.  With blocks and bricks set, decoders
                               without pre-layers use the tile kernel (walk and density per
                               lane, colour in 64-row matrix-core batches; n <= 512) */
    int n;
    int channels;
    float xyzmin;
    float voxel_size;
    const uint32_t* tail;   /* tail bits from diver_tail_bits (of the same mask), or NULL: each
                               walk ends once no occupied voxel can lie ahead on its ray, instead
                               of at the grid's exit (no effect on results) */
} diver_grid;

enum diver_precision { DIVER_FP32 = 0, DIVER_BF16 = 1, DIVER_FP16 = 2 };

enum diver_traversal {
    DIVER_WALK_OFFLINE = 0,  /* masked_intersect's walk from the box entry, exit at > N       */
    DIVER_WALK_REALTIME = 1  /* aabb_intersect's octree descent, then ray_march's walk (>= N) */
};

typedef struct diver_render_opts {
    int precision;     /* diver_precision: decoder arithmetic (traversal/integral stay FP32) */
    int traversal;     /* diver_traversal; n must be a multiple of 64 for the real-time walk  */
    float t_stop;      /* stop a ray once transmittance < t_stop (0 = never)                 */
    float alpha_skip;  /* skip intervals with alpha < alpha_skip, as mlp_eval.cu:153 (0 = off) */
} diver_render_opts;

/* Traversal + integral + decoder + compositing of explicit rays in one launch: the fused form of
 * eval_rt.py's aabb_intersect / ray_march / mlp_eval loop (or of DIVeR.forward + render,
 * model/diver.py:141-207, with the offline walk).  Transmittance is exp(-S) with S a compensated
 * FP32 sum of densities.  rgba [n_rays][4] = (accumulated rgb, T); add T for a white background.
 * count [n_rays] (optional) = intervals composited. */
int diver_render_rays(const diver_grid* grid, const diver_decoder* dec,
                      const diver_render_opts* opts, int64_t n_rays, const float* o,
                      const float* d, float* rgba, int32_t* count, diver_stream_t stream);

/* diver_render_rays plus, per ray, sparsity[r] = sum over the intervals the ray evaluates of
 * log(1 + s^2 / 0.5), s the density pre-activation of the interval: the per-ray part of the
 * reference's sparsity regulariser l_s * mean(log(1 + sigma^2 / 0.5)) over the [B, M] sigma of a
 * training batch (train.py:84-85,110; its zero padding adds nothing).  With t_stop = 0 and
 * alpha_skip = 0 the intervals are all of the ray's occupied intervals, as in the reference's
 * training forward.  Needs the tile path (block and brick bits, a decoder without pre-layers). */
int diver_render_rays_sparsity(const diver_grid* grid, const diver_decoder* dec,
                               const diver_render_opts* opts, int64_t n_rays, const float* o,
                               const float* d, float* rgba, int32_t* count, float* sparsity,
                               diver_stream_t stream);

/* Occupancy extraction (replaces the alpha-map loop of prune.py:83-143: masked_intersect or
 * ray_voxel_intersect, DIVeR.decode, DIVeR.render's weights and torch_scatter's max).  Every
 * occupied voxel a ray crosses (the offline walk over grid->mask; pass an all-ones mask for the
 * unmasked traversal) gets the rendering weight alpha_i * T_i of that interval (no early
 * termination, FP32 decoder); the voxel min(int(p_in + 1e-4), int(p_out + 1e-4)) per axis, if
 * it lies in [pad, N-1-pad]^3, keeps the maximum in alpha_map [N^3] at x + y N + z N^2.  The
 * map is accumulated: zero it once, then call per batch of rays.  Threshold it to get the new
 * voxel mask (prune.py:146). */
int diver_alpha_map(const diver_grid* grid, const diver_decoder* dec, int64_t n_rays,
                    const float* o, const float* d, int pad, float* alpha_map,
                    diver_stream_t stream);

/* Same for a pinhole camera (utils/dataset/ray_utils.py:5-56 conventions: pixel centres +0.5,
 * -z forward, normalised directions); c2w_host is row-major 3x4.  rgba/count are [height][width]. */
int diver_render_camera(const diver_grid* grid, const diver_decoder* dec,
                        const diver_render_opts* opts, const float* c2w_host, int width,
                        int height, float focal, float* rgba, int32_t* count,
                        diver_stream_t stream);

/* One request of diver_render_batch: a pinhole camera (width > 0; c2w row-major 3x4, copied at
 * the call) or n_rays explicit rays o, d [n_rays][3] (width == 0).  Outputs as
 * diver_render_camera / diver_render_rays.  The camera is the Blender one of diver_render_camera
 * (focal, principal point at the image centre) while fy is 0 (a zero-initialised request); with
 * fy > 0 it is the general pinhole camera of diver_camera_rays_pinhole with fx = focal and
 * principal point (cx, cy): the Tanks and Temples / BlendedMVS camera. */
typedef struct diver_request {
    float c2w[12];
    int width;
    int height;
    float focal;
    int64_t n_rays;
    const float* o;
    const float* d;
    float* rgba;
    int32_t* count;
    float fy;
    float cx;
    float cy;
} diver_request;

#define DIVER_MAX_REQUESTS 16

/* Several requests on one grid and decoder in one launch: the rays of all requests are numbered
 * consecutively (a camera's pixels in 8x8 tiles) and every wave renders 64 consecutive ids, so
 * requests of different sizes share the launch.  Each request's pixels are computed exactly as a
 * diver_render_camera / diver_render_rays call of that request alone computes them (a ray's
 * result depends only on its own intervals).  Needs the block and brick bits and a decoder without
 * pre-layers.  Up to DIVER_MAX_REQUESTS requests.
 * No counterpart in the reference (the serving path batches the requests of several clients). */
int diver_render_batch(const diver_grid* grid, const diver_decoder* dec,
                       const diver_render_opts* opts, const diver_request* reqs, int n_req,
                       diver_stream_t stream);

/* Recompute pass of the render adjoint (decoders without pre-layers, alpha_skip == 0).  Given
 * the forward's per-ray interval counts and their exclusive prefix sum row_offset [n_rays]
 * (int64), re-walks every ray and writes, for each composited interval k of ray r, row
 * row_offset[r] + k of rows [sum(count)][9 + channels]:
 *   p_in (3), p_out (3), s = w_s.F + b_s, transmittance before the interval, alpha, F (channels)
 * with the values the forward used.  This replaces the samples the reference's autograd keeps
 * through integrate / mlp / render (model/diver.py:141-207, utils/integrator.py:62-113); the
 * decoder and compositing adjoints run as GEMMs over the rows and diver_integrate_backward
 * scatters the feature gradient into the grid. */
int diver_render_rows(const diver_grid* grid, const diver_decoder* dec,
                      const diver_render_opts* opts, int64_t n_rays, const float* o,
                      const float* d, const int32_t* count, const int64_t* row_offset,
                      float* rows, diver_stream_t stream);

/* The training forward with its intervals recorded: diver_render_rays_sparsity (sparsity may be
 * NULL here) that also writes, for the k-th interval ray r evaluates (k < seg_cap), its entry and
 * exit in grid units to segments [n_rays][seg_cap][6] (p_in, p_out).  FP32 offline walk,
 * alpha_skip == 0, no early termination (t_stop <= 0, as the reference trains), block and brick
 * bits, a decoder without pre-layers whose density is channel 0 (w_s = e_0, b_s = 0).  Without
 * early termination the walk does not need each interval's density, so this kernel queues the
 * intervals and evaluates their densities in the shading batches; rgba, count and sparsity are
 * the same bits as diver_render_rays_sparsity's.  A ray evaluates at most 3 (N + 1) intervals, so seg_cap = 3 (N + 1)
 * records every one.  With diver_render_rows_segments this replaces the re-walk of
 * diver_render_rows, which recomputes the intervals the reference's autograd keeps from
 * masked_intersect (model/diver.py:87-93). */
int diver_render_rays_segments(const diver_grid* grid, const diver_decoder* dec,
                               const diver_render_opts* opts, int64_t n_rays, const float* o,
                               const float* d, float* rgba, int32_t* count, float* sparsity,
                               float* segments, int seg_cap, diver_stream_t stream);

/* diver_render_rows from the segments of diver_render_rays_segments instead of a re-walk: the same
 * rows bit for bit, provided every count <= seg_cap (rows past seg_cap of a ray are not written).
 * ray [sum(count)] (int64, may be NULL) receives the ray of each row. */
int diver_render_rows_segments(const diver_grid* grid, const diver_decoder* dec,
                               const diver_render_opts* opts, int64_t n_rays, const int32_t* count,
                               const int64_t* row_offset, const float* segments, int seg_cap,
                               float* rows, int64_t* ray, diver_stream_t stream);

/* Per-ray suffix sums for the compositing adjoint over the rows of diver_render_rows: for row k of
 * ray r (rows row_offset[r] .. row_offset[r] + count[r] - 1),
 *   suffix[k][ch] = sum over the ray's rows j > k of weight[j] * color[j][ch]  (+ background[r])
 * with FP64 sums; background may be NULL (no background term).  This is the part of the gradient of
 * the front-to-back composite (model/diver.py:196-202, cumprod weights) that torch autograd derives
 * through cumprod in the reference; here dC/dsigma_k = T_after_k c_k - suffix[k]. */
int diver_ray_suffix(int64_t n_rays, const int64_t* row_offset, const int32_t* count,
                     const float* weight, const float* color, const float* background,
                     float* suffix, diver_stream_t stream);

/* Decoder adjoint, forward half, over n_rows rows of diver_render_rows (decoders without
 * pre-layers, one post layer, channels and hidden width multiples of 32; the decoder's current
 * parameters).  ray [n_rows] (int64) names each row's ray, d [R][3] the rays' directions.  Writes,
 * per row, relu(y1) and relu(y2) of the colour MLP (r1, r2 [n_rows][hidden]), the colour
 * color [n_rows][3] and the compositing weight weight [n_rows] = T_before * alpha, evaluated as the
 * fused renderer evaluates them (FP32 operands on the matrix cores).  Together with
 * diver_adjoint_decoder_backward this replaces the autograd of mlp and render over stored samples
 * in the reference's training step (model/diver.py:159-207, train.py:102-113). */
int diver_adjoint_decoder_forward(const diver_decoder* dec, int64_t n_rows, const float* rows,
                                  const int64_t* ray, const float* d, float* r1, float* r2,
                                  float* color, float* weight, diver_stream_t stream);

/* Decoder adjoint, backward half.  Rows n_valid .. n_rows - 1 are padding (zero rows).  With
 * grad_rgb [R][3] = dL/d(rgb), suffix [n_valid][3] from diver_ray_suffix over (weight, color) and
 * sparsity_scale a device scalar (dL/d(sigma_loss) / (B M), or NULL for no sparsity term), writes
 *   d_out [n_rows][3]   dL/d(W_o r2 + b_o) = g w c (1 - c)
 *   d_s [n_rows]        dL/ds = g . (T_after c - suffix) act'(s) (+ sparsity_scale 4s / (1 + 2s^2))
 *   d_r2, d_r1 [n_rows][hidden]   dL/dy2 = (W_o^T d_out) [y2 > 0], dL/dy1 = (W_2^T d_r2) [y1 > 0]
 *   d_feature [n_valid][channels] dL/dF = W_c^T d_r1 + d_s w_s
 * The weight gradients are the reductions d_out^T r2, d_r2^T r1, d_r1^T [F, PE(d)] and d_s^T F over
 * the rows; d_feature goes to diver_integrate_backward. */
int diver_adjoint_decoder_backward(const diver_decoder* dec, int64_t n_rows, int64_t n_valid,
                                   const float* rows, const int64_t* ray, const float* grad_rgb,
                                   const float* suffix, const float* sparsity_scale,
                                   const float* r1, const float* r2, const float* color,
                                   const float* weight, float* d_feature, float* d_out, float* d_s,
                                   float* d_r1, float* d_r2, diver_stream_t stream);

/* Weight gradients of the decoder adjoint over n_rows rows (decoders of diver_adjoint_decoder_*
 * with channels/hidden 32/32, 32/64 or 64/64), from the operands the two halves wrote:
 *   dw_s = d_s^T F, db_s = sum d_s, dW_c = d_r1^T F, dW_v = d_r1^T PE(d[ray]), db_c = sum d_r1,
 *   dW_2 = d_r2^T r1, db_2 = sum d_r2, dW_o = d_out^T r2, db_o = sum d_out
 * (F = rows[:, 9:]), written over grad [diver_decoder_param_count] in the canonical parameter
 * order.  These are the reductions torch autograd forms in the reference's nn.Linear layers of
 * the colour and density MLPs (model/mlps.py:21-31 through model/diver.py:159-171, train.py:110-113).
 * Two launches in a fixed summation order (per-workgroup partials in workspace, then a reduction
 * in workgroup order), so the result does not depend on timing.  workspace: device memory of at
 * least diver_adjoint_weight_grads_workspace(dec, n_rows) bytes; its contents are overwritten. */
int64_t diver_adjoint_weight_grads_workspace(const diver_decoder* dec, int64_t n_rows);
int diver_adjoint_weight_grads(const diver_decoder* dec, int64_t n_rows, const float* rows,
                               const int64_t* ray, const float* d, const float* r1, const float* r2,
                               const float* d_out, const float* d_s, const float* d_r1,
                               const float* d_r2, float* grad, float* workspace,
                               int64_t workspace_bytes, diver_stream_t stream);

/* Allocation-free form of diver_adjoint_weight_grads, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight training loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  The per-workgroup partials live in a file-scope device array
 * of the library (23.1 MB per device, held until the process ends; every decoder shape and row
 * count fits it); the call allocates nothing and otherwise behaves as diver_adjoint_weight_grads,
 * bit for bit.  Every call on one device shares that device's array: calls that overlap on one
 * device (two streams, two threads) race and must be ordered by the caller. */
int diver_adjoint_weight_grads_noalloc(const diver_decoder* dec, int64_t n_rows, const float* rows,
                                       const int64_t* ray, const float* d, const float* r1,
                                       const float* r2, const float* d_out, const float* d_s,
                                       const float* d_r1, const float* d_r2, float* grad,
                                       diver_stream_t stream);

/* The positional encoding PE(d) of n directions, [n][27]: [d, sin(pi d), cos(pi d), ...,
 * sin(8 pi d), cos(8 pi d)] (model/mlps.py:12-18), as the fused kernels evaluate it. */
int diver_positional_encoding(int64_t n, const float* d, float* pe, diver_stream_t stream);

/* 8^3-block occupancy bits of an [n^3] mask (n a multiple of 8): bit b of word b/32 is set
 * when block b = (bz*(n/8) + by)*(n/8) + bx holds an occupied voxel; bits has
 * ceil((n/8)^3 / 32) words.  No counterpart in the reference (an acceleration structure for
 * diver_render_*). */
int diver_block_bits(const uint8_t* mask, int n, uint32_t* bits, diver_stream_t stream);

/* Tail bits of a mask from its 8^3-block bits (diver_block_bits; n a multiple of 8): the block
 * bits dilated by one block -- bit b of word b/32 is set when a block of block b's 3x3x3
 * neighbourhood holds an occupied voxel (same block numbering); tail has ceil((n/8)^3 / 32)
 * words.  A walk from p along d can end once its line has left the last set block it meets:
 * every voxel it could still visit lies within a few voxels of the line, inside an unset block's
 * neighbourhood, so empty.  No counterpart in the reference (an acceleration structure for
 * diver_render_*, diver_alpha_map and diver_walk_rays). */
int diver_tail_bits(const uint32_t* blocks, int n, uint32_t* tail, diver_stream_t stream);

/* Bit-packed mask per 4^3 brick (n a multiple of 4): word b = (bz*(n/4) + by)*(n/4) + bx holds
 * bit (z&3)*16 + (y&3)*4 + (x&3) for voxel (x, y, z) of the brick; bricks has (n/4)^3 words.
 * No counterpart in the reference (an acceleration structure for diver_render_*). */
int diver_brick_bits(const uint8_t* mask, int n, uint64_t* bricks, diver_stream_t stream);

/* One Adam step over n float32 parameters in device memory, replacing the update of
 * torch.optim.Adam(params, lr, weight_decay) that the reference trains with (train.py:39-47):
 *   g' = g + weight_decay * p;  m = beta1 m + (1 - beta1) g';  v = beta2 v + (1 - beta2) g'^2;
 *   p -= step_size * m / (sqrt(v) / bc2_sqrt + eps)
 * with step_size = lr / (1 - beta1^t) and bc2_sqrt = sqrt(1 - beta2^t) for step t, computed by the
 * caller.  The betas are doubles, as torch's are: 1 - beta is rounded to FP32 once from them.
 * clear_grad != 0 also sets the gradient to zero in the same pass.  Arrays that are all 16-byte
 * aligned are processed in float4 groups, others element by element.  One read of p, g, m, v and
 * one write of p, m, v (and g) per parameter. */
int diver_adam_step(float* param, float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                    double beta1, double beta2, float eps, float weight_decay, float step_size,
                    float bc2_sqrt, int clear_grad, diver_stream_t stream);

/* diver_adam_step over listed rows only: row r is elements [r * row_len, (r + 1) * row_len) of
 * the n-element arrays (row_len a multiple of 4 dividing n, arrays 16-byte aligned; rows outside
 * [0, n / row_len) are skipped).  Without weight decay (required: weight_decay == 0, the
 * reference's default, configs/nerf_synthetic_fine.py 'weight_decay') an element whose gradient
 * and both moments are zero is a fixed point of the update, so when every row that ever receives
 * a gradient is listed, this step leaves the arrays exactly as diver_adam_step over all n
 * elements would.  For the trainable grid those rows are the corners of the occupied voxels: the
 * walks composite occupied voxels only, whose integrals read their 8 corners.  No counterpart in
 * the reference (torch.optim.Adam updates the whole grid every step, train.py:39-47). */
int diver_adam_step_rows(float* param, float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                         const int64_t* rows, int64_t n_rows, int row_len, double beta1,
                         double beta2, float eps, float weight_decay, float step_size,
                         float bc2_sqrt, int clear_grad, diver_stream_t stream);

/* The occupied intervals the fused renderer composites, one walk per ray (diver_traversal):
 * the offline walk equals masked_intersect (utils/ray_voxel_intersection/src/
 * ray_voxel_intersect.cu:192-319); the real-time walk equals aabb_intersect followed by ray_march
 * rounds until the ray leaves (utils/ray_march/src/ray_march.cu:24-230), each ray from its own
 * origin.  Same step and occupancy code as diver_render_*; output as diver_masked_intersect
 * (isect [n_rays][max_n][6], num [n_rays]).  grid->blocks is required. */
int diver_walk_rays(const diver_grid* grid, int traversal, int64_t n_rays, const float* o,
                    const float* d, int max_n, float* isect, int32_t* num,
                    diver_stream_t stream);

/* The FP32 rays diver_render_camera traces: o, d [height*width][3]. */
int diver_camera_rays(const float* c2w_host, int width, int height, float focal, float* o,
                      float* d, diver_stream_t stream);

/* The rays of a general pinhole camera: pixel (x, y) looks along
 * ((x + 0.5 - cx)/fx, -(y + 0.5 - cy)/fy, -1) in camera space, rotated by c2w and normalised,
 * from c2w's translation (utils/dataset/tanks.py:120-169, get_ray_directions / get_rays).  The
 * dataset's pose conversion (parse_extrinsics: OpenCV axes to these, tanks.py:172-181), its
 * intrinsics scaling (tanks.py:27-42) and the origin normalisation (o - mid) / scale
 * (tanks.py:46-52, 98) are applied to c2w on the host (diver_amd.cameras). */
int diver_camera_rays_pinhole(const float* c2w_host, int width, int height, float fx, float fy,
                              float cx, float cy, float* o, float* d, diver_stream_t stream);

/* ---- render server ---------------------------------------------------------------------------- */

/* A continuous-batching render loop on one GPU, run by two native threads.  Scenes (a grid plus a
 * decoder and render options) stay resident; clients submit frames (scene, pinhole camera, output
 * buffers) from any thread.  The launch thread packs the pending frames of the oldest frame's
 * scene and size class (FIFO; up to max_requests frames and max_rays ray ids; frames of more than
 * split_rays pixels launch apart from smaller ones) into one diver_render_batch launch, with up to
 * n_streams launches in flight on streams of its own.  Frames that arrive while launches run join
 * the next launch as soon as a stream and memory are free.  The retire thread watches every
 * launch's completion event and completes its frames as soon as the GPU finishes them, in
 * whatever order the launches end.
 * Memory admission: resident scenes (their scene_bytes) and the outputs of the frames on the GPU
 * (20 bytes per pixel: rgba + count) are charged against hbm_budget; a scene or a frame that can
 * never fit is refused with DIVER_EBUDGET, a frame that fits only later waits.
 * A frame's pixels equal diver_render_camera's for the same scene and camera, bit for bit.
 * No counterpart in the reference (its real-time viewer renders one camera at a time,
 * eval_rt.py:106-135); this is the serving path around the same renderer. */
typedef struct diver_server diver_server;

typedef struct diver_server_config {
    int64_t hbm_budget;  /* bytes for resident scenes plus the outputs of frames on the GPU     */
    int n_streams;       /* launches in flight, 1..8                                           */
    int max_requests;    /* frames per launch, 1..DIVER_MAX_REQUESTS                            */
    int64_t max_rays;    /* ray ids per launch (a camera takes 64 per 8x8 tile)                 */
    int64_t split_rays;  /* frames of more pixels launch apart from smaller ones; 0 mixes them  */
} diver_server_config;

typedef struct diver_frame_info {
    int64_t ticket;
    int scene;
    int width;
    int height;
    int status;          /* DIVER_OK once rendered; an error code if its launch failed          */
    int64_t launch;      /* serial number of the launch that rendered it (-1 before launch)    */
    double t_submit;     /* diver_server_clock() at submit, at launch, at completion (seconds)  */
    double t_launch;
    double t_done;
} diver_frame_info;

/* Seconds on the monotonic clock the server stamps frames with (CLOCK_MONOTONIC). */
double diver_server_clock(void);
/* Creates the server on the calling thread's current device and starts its two threads. */
int diver_server_create(const diver_server_config* cfg, diver_server** out);
/* Renders every frame already submitted, then stops the threads and frees the server. */
int diver_server_destroy(diver_server* srv);
/* Makes a scene resident (the grid's and decoder's device memory stay the caller's and must
 * outlive the server).  Work the caller enqueued on `ready` before this call (building the grid)
 * is ordered before every launch of the scene.  Needs the block and brick bits and a decoder
 * without pre-layers.  *scene_id receives the scene's number. */
int diver_server_add_scene(diver_server* srv, const diver_grid* grid, const diver_decoder* dec,
                           const diver_render_opts* opts, int64_t scene_bytes,
                           diver_stream_t ready, int* scene_id);
/* Queues one frame: c2w_host row-major 3x4 (copied), rgba [height][width][4] and count
 * [height][width] (optional) device buffers that the caller keeps alive until the frame is
 * complete.  Work the caller enqueued on `after` before this call (e.g. earlier use of the
 * output memory) is ordered before the frame's launch; NULL is the null stream.  Returns at once
 * with *ticket. */
int diver_server_submit(diver_server* srv, int scene_id, const float* c2w_host, int width,
                        int height, float focal, float* rgba, int32_t* count,
                        diver_stream_t after, int64_t* ticket);
/* The same for a general pinhole camera (fx, fy, principal point (cx, cy); see
 * diver_camera_rays_pinhole): Tanks and Temples / BlendedMVS views. */
int diver_server_submit_pinhole(diver_server* srv, int scene_id, const float* c2w_host, int width,
                                int height, float fx, float fy, float cx, float cy, float* rgba,
                                int32_t* count, diver_stream_t after, int64_t* ticket);
/* Blocks until the frame is complete (or timeout_s seconds pass: DIVER_ETIMEDOUT; timeout_s < 0
 * waits without limit) and fills *info (may be NULL).  A completed frame is waited for at most
 * once. */
int diver_server_wait(diver_server* srv, int64_t ticket, double timeout_s, diver_frame_info* info);
/* Tickets of the frames completed since the previous call (at most max), oldest first; returns
 * how many were written. */
int64_t diver_server_take_completed(diver_server* srv, int64_t* tickets, int64_t max);
/* The records of the last completed frames (at most max, oldest first, up to 65536 kept) and the
 * number of launches so far; returns how many records were written. */
int64_t diver_server_history(diver_server* srv, diver_frame_info* out, int64_t max,
                             int64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* DIVER_H */
