/*
 * CPU ORACLE -- test infrastructure only.  Nothing in the product path may link or call this
 * file; only tests/, the smoke check and bench.py's CPU-baseline leg use it, as the checker.
 *
 * FP32 restatement of the reference's voxel traversal kernels, one ray after another, in the
 * reference's own operation order:
 *   oracle_ray_voxel_intersect  <- utils/ray_voxel_intersection/src/ray_voxel_intersect.cu:29-151
 *   oracle_masked_intersect     <- utils/ray_voxel_intersection/src/ray_voxel_intersect.cu:192-319
 *   oracle_aabb_intersect       <- utils/ray_march/src/ray_march.cu:24-149
 *   oracle_ray_march            <- utils/ray_march/src/ray_march.cu:151-230
 *
 * Contraction policy (stated, because the recurrence is bit-sensitive): the reference is built by
 * nvcc with its default --fmad=true, which fuses a multiply feeding an add.  The sites where that
 * happens are written here as explicit fmaf():  `o + d*tmin` (box entry) and `p + d*tnext` (DDA
 * step).  `floor(p*step + 1)` is exact with or without fusion because step is -1, 0 or 1.  Every
 * other operation is a single IEEE round-to-nearest op; this file is compiled with
 * -ffp-contract=off so the C compiler adds no fusion of its own.  Divisions are IEEE divisions.
 *
 * Two deliberate, documented deviations from the reference, both on inputs where the reference
 * has undefined behaviour:
 *   - every DDA loop is bounded (max_steps below); the reference loops forever on a zero or NaN
 *     direction;
 *   - writes past max_n entries are dropped (the count still advances); the reference writes past
 *     the end of the ray's row.
 * Octree/mask reads outside the array are treated as empty (the reference reads out of bounds).
 *
 * parity: pinned by tests/test_oracle_traversal.py (geometric invariants, brute-force voxel
 * sequences on small grids, axis-aligned and edge/corner cases).
 */
#include <math.h>
#include <stdint.h>

static inline int sgn(float x) {
    int t = x < 0 ? -1 : 0;
    return x > 0 ? 1 : t;
}

/* cutil-style clamp: fmaxf(a, fminf(f, b)) */
static inline float clampf_(float f, float a, float b) { return fmaxf(a, fminf(f, b)); }

static inline int max_steps_for(float voxel_num) {
    float n = voxel_num > 0 ? voxel_num : 0;
    if (!(n < 1e7f)) n = 1e7f;
    return 3 * (int)ceilf(n) + 8;
}

/* slab test against [0, n]^3 in grid units; returns 1 on hit; p is clamped entry point.
 * Operation order follows ray_voxel_intersect.cu:234-251. */
static int box_entry(float ox, float oy, float oz, float dx, float dy, float dz, float n,
                     float *px, float *py, float *pz, float *tentry) {
    float t0 = (-ox) / dx;
    float t1 = (n - ox) / dx;
    float tmin = fminf(t0, t1);
    float tmax = fmaxf(t0, t1);
    t0 = (-oy) / dy;
    t1 = (n - oy) / dy;
    tmin = fmaxf(tmin, fminf(t0, t1));
    tmax = fminf(tmax, fmaxf(t0, t1));
    t0 = (-oz) / dz;
    t1 = (n - oz) / dz;
    tmin = fmaxf(tmin, fminf(t0, t1));
    tmax = fminf(tmax, fmaxf(t0, t1));
    *px = clampf_(fmaf(dx, tmin, ox), 0.0f, n);
    *py = clampf_(fmaf(dy, tmin, oy), 0.0f, n);
    *pz = clampf_(fmaf(dz, tmin, oz), 0.0f, n);
    *tentry = tmin;
    return !(tmin > tmax);
}

/* one DDA step from (px,py,pz): next boundary crossing, snapped; ray_voxel_intersect.cu:272-288 */
static inline float dda_step(float px, float py, float pz, float dx, float dy, float dz,
                             float sx, float sy, float sz, float *qx, float *qy, float *qz) {
    float bx = floorf(px * sx + 1.0f) * sx;
    float by = floorf(py * sy + 1.0f) * sy;
    float bz = floorf(pz * sz + 1.0f) * sz;
    float tx = (bx - px) / dx;
    float ty = (by - py) / dy;
    float tz = (bz - pz) / dz;
    float tn = fminf(tx, fminf(ty, tz));
    float nx = fmaf(dx, tn, px);
    float ny = fmaf(dy, tn, py);
    float nz = fmaf(dz, tn, pz);
    if (tn == tx) {
        nx = bx;
    } else if (tn == ty) {
        ny = by;
    } else {
        nz = bz;
    }
    *qx = nx;
    *qy = ny;
    *qz = nz;
    return tn;
}

void oracle_ray_voxel_intersect(int64_t batch, int max_n, float xyzmin, float xyzmax,
                                float voxel_num, float voxel_size, const float *o, const float *v,
                                float *isect, int32_t *num, float *tns) {
    (void)xyzmax;
    int max_steps = max_steps_for(voxel_num);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < batch; ++b) {
        const float *ob = o + 3 * b;
        const float *vb = v + 3 * b;
        float *ib = isect + (int64_t)3 * max_n * b;
        float *tb = tns + (int64_t)max_n * b;
        int32_t count = 0;
        float ox = (ob[0] - xyzmin) / voxel_size;
        float oy = (ob[1] - xyzmin) / voxel_size;
        float oz = (ob[2] - xyzmin) / voxel_size;
        float dx = vb[0], dy = vb[1], dz = vb[2];
        float px = ox, py = oy, pz = oz;
        int inside = (ox >= 0) & (oy >= 0) & (oz >= 0) & (ox <= voxel_num) & (oy <= voxel_num) &
                     (oz <= voxel_num);
        float t_now;
        if (inside) {
            if (max_n > 0) {
                ib[0] = px; ib[1] = py; ib[2] = pz;
                tb[0] = 0.0f;
            }
            t_now = 0.0f;
        } else {
            float tmin;
            int hit = box_entry(ox, oy, oz, dx, dy, dz, voxel_num, &px, &py, &pz, &tmin);
            if (max_n > 0) {
                ib[0] = px; ib[1] = py; ib[2] = pz;
            }
            if (!hit) {
                num[b] = 0;
                continue;
            }
            if (max_n > 0) tb[0] = tmin;
            t_now = tmin;
        }
        if (max_n > 0) tb[0] = t_now * voxel_size;
        count = 1;
        float sx = (float)sgn(dx), sy = (float)sgn(dy), sz = (float)sgn(dz);
        for (int it = 0; it < max_steps; ++it) {
            float qx, qy, qz;
            float tn = dda_step(px, py, pz, dx, dy, dz, sx, sy, sz, &qx, &qy, &qz);
            t_now += tn;
            px = qx; py = qy; pz = qz;
            if ((px < 0) | (py < 0) | (pz < 0) | (px > voxel_num) | (py > voxel_num) |
                (pz > voxel_num))
                break;
            if (count < max_n) {
                ib[3 * count + 0] = px;
                ib[3 * count + 1] = py;
                ib[3 * count + 2] = pz;
                tb[count] = t_now * voxel_size;
            }
            count += 1;
        }
        num[b] = count;
    }
}

void oracle_masked_intersect(int64_t batch, int max_n, float xyzmin, float xyzmax,
                             float voxel_num, float voxel_size, float mask_scale, const float *o,
                             const float *v, const uint8_t *mask, float *isect, int32_t *num,
                             float *tns) {
    (void)xyzmax;
    int max_steps = max_steps_for(voxel_num);
    int mask_size = (int)(voxel_num * mask_scale);
    int64_t mask_total = (int64_t)mask_size * mask_size * mask_size;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < batch; ++b) {
        const float *ob = o + 3 * b;
        const float *vb = v + 3 * b;
        float *ib = isect ? isect + (int64_t)6 * max_n * b : 0;
        float *tb = tns ? tns + (int64_t)max_n * b : 0;
        int32_t count = 0;
        float ox = (ob[0] - xyzmin) / voxel_size;
        float oy = (ob[1] - xyzmin) / voxel_size;
        float oz = (ob[2] - xyzmin) / voxel_size;
        float dx = vb[0], dy = vb[1], dz = vb[2];
        float lx = ox, ly = oy, lz = oz; /* ori_last */
        int inside = (ox >= 0) & (oy >= 0) & (oz >= 0) & (ox <= voxel_num) & (oy <= voxel_num) &
                     (oz <= voxel_num);
        float t_now;
        if (inside) {
            t_now = 0.0f;
        } else {
            float tmin;
            int hit = box_entry(ox, oy, oz, dx, dy, dz, voxel_num, &lx, &ly, &lz, &tmin);
            if (!hit) {
                num[b] = 0;
                continue;
            }
            t_now = tmin;
        }
        float sx = (float)sgn(dx), sy = (float)sgn(dy), sz = (float)sgn(dz);
        for (int it = 0; it < max_steps; ++it) {
            float px, py, pz;
            float tn = dda_step(lx, ly, lz, dx, dy, dz, sx, sy, sz, &px, &py, &pz);
            t_now += tn;
            if ((px < 0) | (py < 0) | (pz < 0) | (px > voxel_num) | (py > voxel_num) |
                (pz > voxel_num))
                break;
            float cx = fminf(lx + 1e-4f, px + 1e-4f) * mask_scale;
            float cy = fminf(ly + 1e-4f, py + 1e-4f) * mask_scale;
            float cz = fminf(lz + 1e-4f, pz + 1e-4f) * mask_scale;
            int64_t ci = (int64_t)(int)cz * mask_size * mask_size + (int64_t)(int)cy * mask_size +
                         (int)cx;
            int occ = (ci >= 0 && ci < mask_total) ? mask[ci] != 0 : 0;
            if (occ) {
                if (count < max_n) {
                    if (ib) {
                        float *r = ib + 6 * count;
                        r[0] = lx; r[1] = ly; r[2] = lz;
                        r[3] = px; r[4] = py; r[5] = pz;
                    }
                    if (tb) tb[count] = t_now * voxel_size;
                }
                count += 1;
            }
            lx = px; ly = py; lz = pz;
        }
        num[b] = count;
    }
}

#define MAX_HIT 8

void oracle_aabb_intersect(int64_t batch, float *coords, const float *v, const float *center,
                           uint8_t *finish, float *rgba, const uint8_t *octrees, int voxel_num) {
    float n = (float)voxel_num;
    int64_t oct_total = 0;
    for (int m = voxel_num / 64; m < voxel_num; m *= 4) oct_total += (int64_t)m * m * m;
    int max_steps = max_steps_for(n) + 16;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < batch; ++b) {
        float *o = coords + batch * 6 * (MAX_HIT - 1) + b * 6 + 3;
        const float *vb = v + 3 * b;
        rgba[4 * b + 0] = 0.0f;
        rgba[4 * b + 1] = 0.0f;
        rgba[4 * b + 2] = 0.0f;
        rgba[4 * b + 3] = 1.0f;
        finish[b] = 0;
        float px = center[0], py = center[1], pz = center[2];
        float dx = vb[0], dy = vb[1], dz = vb[2];
        int inside = (px >= 0) & (py >= 0) & (pz >= 0) & (px <= n) & (py <= n) & (pz <= n);
        if (!inside) {
            float tmin;
            int hit = box_entry(px, py, pz, dx, dy, dz, n, &px, &py, &pz, &tmin);
            if (!hit) {
                o[0] = -1.0f;
                finish[b] = 1;
                continue;
            }
        }
        px /= 64.0f; py /= 64.0f; pz /= 64.0f;
        float lx = px, ly = py, lz = pz;
        float sx = (float)sgn(dx), sy = (float)sgn(dy), sz = (float)sgn(dz);
        int mask_size = voxel_num / 64;
        int64_t off = 0;
        int done = 0;
        for (int it = 0; it < max_steps && mask_size < voxel_num; ++it) {
            lx = px; ly = py; lz = pz;
            float tn = dda_step(lx, ly, lz, dx, dy, dz, sx, sy, sz, &px, &py, &pz);
            (void)tn;
            float ms = (float)mask_size;
            if ((px < 0) | (py < 0) | (pz < 0) | (px > ms) | (py > ms) | (pz > ms)) {
                o[0] = -1.0f;
                finish[b] = 1;
                done = 1;
                break;
            }
            float cx = fminf(lx + 1e-4f, px + 1e-4f);
            float cy = fminf(ly + 1e-4f, py + 1e-4f);
            float cz = fminf(lz + 1e-4f, pz + 1e-4f);
            int64_t ci = (int64_t)(int)cz * mask_size * mask_size + (int64_t)(int)cy * mask_size +
                         (int)cx;
            int64_t gi = off + ci;
            int occ = (ci >= 0 && ci < (int64_t)mask_size * mask_size * mask_size && gi < oct_total)
                          ? octrees[gi] != 0
                          : 0;
            if (occ) {
                off += (int64_t)mask_size * mask_size * mask_size;
                lx *= 4.0f; ly *= 4.0f; lz *= 4.0f;
                px = lx; py = ly; pz = lz;
                mask_size *= 4;
            }
        }
        if (done) continue;
        if (mask_size < voxel_num) { /* step bound hit: treat as a miss */
            o[0] = -1.0f;
            finish[b] = 1;
            continue;
        }
        o[0] = lx;
        o[1] = ly;
        o[2] = lz;
    }
}

/* Whole real-time walk of each ray: aabb_intersect's start point (its own origin per ray, grid
 * units (o - xyzmin)/voxel_size as eval_rt.py:110 computes the shared centre) followed by
 * ray_march rounds until the ray leaves at >= N.  Rounds do not change the recurrence (a round
 * resumes from the exit of the previous round's last interval), so this emits exactly the
 * intervals that aabb_intersect + repeated ray_march produce.  Output as masked_intersect. */
void oracle_rt_walk(int64_t batch, int max_n, float xyzmin, float voxel_size, int voxel_num,
                    const float *o, const float *v, const uint8_t *octrees, const uint8_t *mask,
                    float *isect, int32_t *num) {
    float n = (float)voxel_num;
    int64_t oct_total = 0;
    for (int m = voxel_num / 64; m < voxel_num; m *= 4) oct_total += (int64_t)m * m * m;
    int64_t total = (int64_t)voxel_num * voxel_num * voxel_num;
    int max_steps = max_steps_for(n);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < batch; ++b) {
        num[b] = 0;
        float px = (o[3 * b] - xyzmin) / voxel_size;
        float py = (o[3 * b + 1] - xyzmin) / voxel_size;
        float pz = (o[3 * b + 2] - xyzmin) / voxel_size;
        float dx = v[3 * b], dy = v[3 * b + 1], dz = v[3 * b + 2];
        int inside = (px >= 0) & (py >= 0) & (pz >= 0) & (px <= n) & (py <= n) & (pz <= n);
        if (!inside) {
            float tmin;
            if (!box_entry(px, py, pz, dx, dy, dz, n, &px, &py, &pz, &tmin)) continue;
        }
        px /= 64.0f; py /= 64.0f; pz /= 64.0f;
        float lx = px, ly = py, lz = pz;
        float sx = (float)sgn(dx), sy = (float)sgn(dy), sz = (float)sgn(dz);
        int m = voxel_num / 64;
        int64_t off = 0;
        int dead = 0;
        for (int it = 0; it < max_steps + 16 && m < voxel_num; ++it) {
            lx = px; ly = py; lz = pz;
            dda_step(lx, ly, lz, dx, dy, dz, sx, sy, sz, &px, &py, &pz);
            float ms = (float)m;
            if ((px < 0) | (py < 0) | (pz < 0) | (px > ms) | (py > ms) | (pz > ms)) {
                dead = 1;
                break;
            }
            float cx = fminf(lx + 1e-4f, px + 1e-4f);
            float cy = fminf(ly + 1e-4f, py + 1e-4f);
            float cz = fminf(lz + 1e-4f, pz + 1e-4f);
            int64_t m3 = (int64_t)m * m * m;
            int64_t ci = (int64_t)(int)cz * m * m + (int64_t)(int)cy * m + (int)cx;
            int64_t gi = off + ci;
            int occ = (ci >= 0 && ci < m3 && gi < oct_total) ? octrees[gi] != 0 : 0;
            if (occ) {
                off += m3;
                lx *= 4.0f; ly *= 4.0f; lz *= 4.0f;
                px = lx; py = ly; pz = lz;
                m *= 4;
            }
        }
        if (dead || m < voxel_num) continue;
        int32_t count = 0;
        for (int it = 0; it < max_steps; ++it) {
            float qx, qy, qz;
            dda_step(lx, ly, lz, dx, dy, dz, sx, sy, sz, &qx, &qy, &qz);
            if ((qx < 0) | (qy < 0) | (qz < 0) | (qx >= n) | (qy >= n) | (qz >= n)) break;
            float cx = fminf(lx + 1e-4f, qx + 1e-4f);
            float cy = fminf(ly + 1e-4f, qy + 1e-4f);
            float cz = fminf(lz + 1e-4f, qz + 1e-4f);
            int64_t ci = (int64_t)(int)cz * voxel_num * voxel_num + (int64_t)(int)cy * voxel_num +
                         (int)cx;
            int occ = (ci >= 0 && ci < total) ? mask[ci] != 0 : 0;
            if (occ) {
                if (count < max_n) {
                    float *r = isect + ((int64_t)b * max_n + count) * 6;
                    r[0] = lx; r[1] = ly; r[2] = lz;
                    r[3] = qx; r[4] = qy; r[5] = qz;
                }
                count += 1;
            }
            lx = qx; ly = qy; lz = qz;
        }
        num[b] = count;
    }
}

/* One round of the real-time fine march.  `stride` reproduces the reference's grid-stride
 * assignment (NUM_BLOCKS*NUM_THREADS = 262144 rays per thread sweep): a thread that meets a
 * resumable ray with a negative resume point leaves its loop, so the later rays of that thread
 * are not marched in this round (ray_march.cu:161-173). */
void oracle_ray_march(int64_t batch, float *coords, const float *v, const uint8_t *mask,
                      const uint8_t *finish, int voxel_num, int64_t stride) {
    float n = (float)voxel_num;
    int64_t total = (int64_t)voxel_num * voxel_num * voxel_num;
    int max_steps = max_steps_for(n);
    if (stride <= 0) stride = batch;
    int64_t lim = stride < batch ? stride : batch;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t t0 = 0; t0 < lim; ++t0) {
        for (int64_t b = t0; b < batch; b += stride) {
            if (finish[b]) continue;
            int64_t last = (int64_t)(MAX_HIT - 1) * 6 * batch + b * 6 + 3;
            float lx = coords[last], ly = coords[last + 1], lz = coords[last + 2];
            coords[b * 6] = -1.0f;
            if (lx < 0.0f) break;
            float dx = v[3 * b], dy = v[3 * b + 1], dz = v[3 * b + 2];
            float sx = (float)sgn(dx), sy = (float)sgn(dy), sz = (float)sgn(dz);
            int hit_num = 0;
            for (int it = 0; it < max_steps && hit_num < MAX_HIT; ++it) {
                float px, py, pz;
                dda_step(lx, ly, lz, dx, dy, dz, sx, sy, sz, &px, &py, &pz);
                if ((px < 0) | (py < 0) | (pz < 0) | (px >= n) | (py >= n) | (pz >= n)) {
                    coords[b * 6 + (int64_t)hit_num * 6 * batch] = -1.0f;
                    break;
                }
                float cx = fminf(lx + 1e-4f, px + 1e-4f);
                float cy = fminf(ly + 1e-4f, py + 1e-4f);
                float cz = fminf(lz + 1e-4f, pz + 1e-4f);
                int64_t ci = (int64_t)(int)cz * voxel_num * voxel_num +
                             (int64_t)(int)cy * voxel_num + (int)cx;
                int occ = (ci >= 0 && ci < total) ? mask[ci] != 0 : 0;
                if (occ) {
                    int64_t r = b * 6 + (int64_t)hit_num * 6 * batch;
                    coords[r] = lx; coords[r + 1] = ly; coords[r + 2] = lz;
                    coords[r + 3] = px; coords[r + 4] = py; coords[r + 5] = pz;
                    hit_num += 1;
                }
                lx = px; ly = py; lz = pz;
            }
        }
    }
}
