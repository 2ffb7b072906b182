// Continuous-batching render server (include/diver.h, diver_server_*): a launch thread packs
// pending frames into diver_render_batch launches on the server's own streams, a retire thread
// completes them as their events fire.  All state is guarded by one mutex; the threads never hold
// it across a HIP call that can block.
#include <hip/hip_runtime.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/diver.h"

namespace diver {
int set_error(int code, const std::string& msg);
}

namespace {

constexpr int kMaxStreams = 8;
constexpr int64_t kBytesPerPixel = 20;  // rgba fp32 + int32 count
constexpr size_t kHistory = 65536;
constexpr size_t kKeepFinished = 1 << 20;
constexpr double kRefresh = 50e-6;     // the retire thread re-reads the launch list this often
constexpr size_t kSpareEvents = 4096;  // completed frames' ordering events kept for reuse

double now_s() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int err(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int err(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return diver::set_error(code, buf);
}

struct Scene {
    diver_grid grid;
    const diver_decoder* dec;
    diver_render_opts opts;
    int64_t bytes;
    hipEvent_t ready;
};

struct Frame {
    diver_frame_info info;
    float c2w[12];
    float focal;
    float fy, cx, cy;  // a general pinhole camera when fy > 0 (diver_request's fields)
    float* rgba;
    int32_t* count;
    int64_t bytes;
    int64_t ray_ids;
    bool large;
    hipEvent_t after;  // recorded on the caller's stream at submit
};

struct Launch {
    int slot;
    std::vector<Frame*> frames;
    int64_t bytes;
};

}  // namespace

struct diver_server {
    diver_server_config cfg{};
    int device = 0;
    hipStream_t streams[kMaxStreams] = {};
    hipEvent_t events[kMaxStreams] = {};
    bool slot_busy[kMaxStreams] = {};
    int next_slot = 0;
    std::vector<Scene> scenes;
    int64_t resident = 0;
    int64_t inflight_bytes = 0;
    std::deque<Frame*> pending;
    std::vector<hipEvent_t> spare_events;                     // reused by submit
    std::vector<Launch> inflight;
    std::unordered_map<int64_t, Frame*> active;               // submitted, not complete
    std::unordered_map<int64_t, diver_frame_info> finished;   // complete, not waited for
    std::deque<int64_t> finished_order;
    std::deque<int64_t> completed;                            // for take_completed
    std::vector<diver_frame_info> history;
    size_t hist_next = 0, hist_count = 0;
    int64_t next_ticket = 1;
    int64_t launches = 0;
    bool stopping = false;
    bool launcher_done = false;
    bool failed = false;
    std::string fail_msg;
    std::mutex mu;
    std::condition_variable cv_launch, cv_retire, cv_done;
    std::thread launcher, retirer;
};

namespace {

// Keeps a frame's ordering event for a later submit (the event has fired or was never waited on:
// a launch waits on it before the frame completes).  Lock held.
void release_event(diver_server* s, hipEvent_t e) {
    if (!e) return;
    if (s->spare_events.size() < kSpareEvents)
        s->spare_events.push_back(e);
    else
        (void)hipEventDestroy(e);
}

// Moves a frame to the finished state.  Lock held.
void finish_frame(diver_server* s, Frame* f, int status, double t) {
    f->info.status = status;
    f->info.t_done = t;
    s->active.erase(f->info.ticket);
    s->finished[f->info.ticket] = f->info;
    s->finished_order.push_back(f->info.ticket);
    while (s->finished_order.size() > kKeepFinished) {
        s->finished.erase(s->finished_order.front());
        s->finished_order.pop_front();
    }
    s->completed.push_back(f->info.ticket);
    if (s->completed.size() > kKeepFinished) s->completed.pop_front();
    s->history[s->hist_next] = f->info;
    s->hist_next = (s->hist_next + 1) % kHistory;
    s->hist_count = std::min(s->hist_count + 1, kHistory);
    release_event(s, f->after);
    delete f;
}

bool size_class(const diver_server*, const Frame* f) { return f->large; }

// The pending frames of the oldest frame's scene and size class that fit the launch and the
// memory budget, FIFO within them.  Lock held.
std::vector<Frame*> take_batch(diver_server* s) {
    std::vector<Frame*> batch;
    if (s->pending.empty()) return batch;
    const int scene = s->pending.front()->info.scene;
    const bool cls = size_class(s, s->pending.front());
    const int64_t free_bytes = s->cfg.hbm_budget - s->resident - s->inflight_bytes;
    int64_t rays = 0, bytes = 0;
    for (Frame* f : s->pending) {
        if (f->info.scene != scene || size_class(s, f) != cls) continue;
        if ((int)batch.size() == s->cfg.max_requests || rays + f->ray_ids > s->cfg.max_rays ||
            bytes + f->bytes > free_bytes)
            break;
        batch.push_back(f);
        rays += f->ray_ids;
        bytes += f->bytes;
    }
    for (Frame* f : batch) s->pending.erase(std::find(s->pending.begin(), s->pending.end(), f));
    return batch;
}

void launch_loop(diver_server* s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        std::vector<Frame*> batch;
        s->cv_launch.wait(lk, [&] {
            if (s->failed || (s->stopping && s->pending.empty())) return true;
            if (s->pending.empty() || (int)s->inflight.size() >= s->cfg.n_streams) return false;
            batch = take_batch(s);
            return !batch.empty();
        });
        if (s->failed) {
            const double t = now_s();
            for (Frame* f : s->pending) finish_frame(s, f, DIVER_ESTATE, t);
            s->pending.clear();
            break;
        }
        if (batch.empty()) break;  // stopping with nothing pending
        int slot = -1;
        for (int k = 0; k < s->cfg.n_streams; ++k) {
            const int c = (s->next_slot + k) % s->cfg.n_streams;
            if (!s->slot_busy[c]) {
                slot = c;
                break;
            }
        }
        s->next_slot = (slot + 1) % s->cfg.n_streams;
        s->slot_busy[slot] = true;
        int64_t bytes = 0;
        for (Frame* f : batch) bytes += f->bytes;
        s->inflight_bytes += bytes;
        const int64_t serial = s->launches++;
        const Scene sc = s->scenes[batch[0]->info.scene];
        const double t = now_s();
        for (Frame* f : batch) {
            f->info.t_launch = t;
            f->info.launch = serial;
        }
        lk.unlock();
        // outside the lock: enqueue the launch and its completion event
        diver_request reqs[DIVER_MAX_REQUESTS];
        std::memset(reqs, 0, sizeof(reqs));
        hipStream_t st = s->streams[slot];
        hipError_t he = hipStreamWaitEvent(st, sc.ready, 0);
        for (size_t i = 0; i < batch.size(); ++i) {
            const Frame* f = batch[i];
            std::memcpy(reqs[i].c2w, f->c2w, sizeof(f->c2w));
            reqs[i].width = f->info.width;
            reqs[i].height = f->info.height;
            reqs[i].focal = f->focal;
            reqs[i].fy = f->fy;
            reqs[i].cx = f->cx;
            reqs[i].cy = f->cy;
            reqs[i].rgba = f->rgba;
            reqs[i].count = f->count;
            if (he == hipSuccess) he = hipStreamWaitEvent(st, f->after, 0);
        }
        int status = he == hipSuccess ? DIVER_OK : DIVER_EHIP;
        std::string msg = he == hipSuccess ? "" : hipGetErrorString(he);
        if (status == DIVER_OK) {
            status = diver_render_batch(&sc.grid, sc.dec, &sc.opts, reqs, (int)batch.size(),
                                        reinterpret_cast<diver_stream_t>(st));
            if (status != DIVER_OK) msg = diver_last_error();
        }
        if (status == DIVER_OK) {
            he = hipEventRecord(s->events[slot], st);
            if (he != hipSuccess) {
                status = DIVER_EHIP;
                msg = hipGetErrorString(he);
            }
        }
        lk.lock();
        if (status != DIVER_OK) {
            // the launch was refused: its frames fail, the server goes on
            const double td = now_s();
            for (Frame* f : batch) finish_frame(s, f, status, td);
            s->inflight_bytes -= bytes;
            s->slot_busy[slot] = false;
            s->fail_msg = msg;
            s->cv_done.notify_all();
            continue;
        }
        s->inflight.push_back(Launch{slot, std::move(batch), bytes});
        s->cv_retire.notify_one();
    }
    s->launcher_done = true;
    s->cv_retire.notify_all();
    s->cv_done.notify_all();
}

// Polls the completion events of every launch on the GPU (yielding between polls, so a launch
// that ends early on one stream is retired without waiting for an older one on another) and
// completes the frames of each launch as soon as its event fires.
void retire_loop(diver_server* s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        s->cv_retire.wait(lk, [&] { return !s->inflight.empty() || s->launcher_done; });
        if (s->inflight.empty()) break;  // the launcher is done and nothing is on the GPU
        std::vector<std::pair<int, hipEvent_t>> watch;
        for (const Launch& l : s->inflight) watch.emplace_back(l.slot, s->events[l.slot]);
        lk.unlock();
        std::vector<std::pair<int, hipError_t>> ended;
        const double t0 = now_s();
        while (true) {
            for (const auto& w : watch) {
                const hipError_t q = hipEventQuery(w.second);
                if (q != hipErrorNotReady) ended.emplace_back(w.first, q);
            }
            if (!ended.empty() || now_s() - t0 > kRefresh) break;
            std::this_thread::yield();
        }
        const double t = now_s();
        lk.lock();
        if (ended.empty()) continue;  // pick up launches queued meanwhile
        for (const auto& e : ended) {
            auto it = std::find_if(s->inflight.begin(), s->inflight.end(),
                                   [&](const Launch& l) { return l.slot == e.first; });
            if (it == s->inflight.end()) continue;
            int status = DIVER_OK;
            if (e.second != hipSuccess) {
                status = DIVER_EHIP;
                s->failed = true;
                s->fail_msg = std::string("launch failed: ") + hipGetErrorString(e.second);
            }
            for (Frame* f : it->frames) finish_frame(s, f, status, t);
            s->inflight_bytes -= it->bytes;
            s->slot_busy[it->slot] = false;
            s->inflight.erase(it);
        }
        s->cv_done.notify_all();
        s->cv_launch.notify_all();
    }
}

}  // namespace

extern "C" {

double diver_server_clock(void) { return now_s(); }

int diver_server_create(const diver_server_config* cfg, diver_server** out) {
    if (!cfg || !out) return err(DIVER_EINVAL, "null argument");
    if (cfg->hbm_budget <= 0) return err(DIVER_EINVAL, "hbm_budget must be positive");
    if (cfg->n_streams < 1 || cfg->n_streams > kMaxStreams)
        return err(DIVER_EINVAL, "n_streams must be in [1, %d]", kMaxStreams);
    if (cfg->max_requests < 1 || cfg->max_requests > DIVER_MAX_REQUESTS)
        return err(DIVER_EINVAL, "max_requests must be in [1, %d]", DIVER_MAX_REQUESTS);
    if (cfg->max_rays < 64) return err(DIVER_EINVAL, "max_rays must be at least 64");
    if (cfg->split_rays < 0) return err(DIVER_EINVAL, "split_rays < 0");
    auto* s = new diver_server();
    s->cfg = *cfg;
    s->history.resize(kHistory);
    hipError_t he = hipGetDevice(&s->device);
    for (int k = 0; k < cfg->n_streams && he == hipSuccess; ++k) {
        // default priority: streams of other priorities are not used anywhere in this library
        he = hipStreamCreateWithFlags(&s->streams[k], hipStreamNonBlocking);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&s->events[k], hipEventDisableTiming);
    }
    if (he != hipSuccess) {
        for (int k = 0; k < kMaxStreams; ++k) {
            if (s->events[k]) (void)hipEventDestroy(s->events[k]);
            if (s->streams[k]) (void)hipStreamDestroy(s->streams[k]);
        }
        delete s;
        return err(DIVER_EHIP, "server streams: %s", hipGetErrorString(he));
    }
    s->launcher = std::thread(launch_loop, s);
    s->retirer = std::thread(retire_loop, s);
    *out = s;
    return DIVER_OK;
}

int diver_server_destroy(diver_server* s) {
    if (!s) return DIVER_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->cv_launch.notify_all();
        s->cv_retire.notify_all();
    }
    s->launcher.join();
    s->retirer.join();
    (void)hipSetDevice(s->device);
    for (int k = 0; k < s->cfg.n_streams; ++k) {
        (void)hipEventDestroy(s->events[k]);
        (void)hipStreamDestroy(s->streams[k]);
    }
    for (Scene& sc : s->scenes) (void)hipEventDestroy(sc.ready);
    for (auto& kv : s->active) {
        if (kv.second->after) (void)hipEventDestroy(kv.second->after);
        delete kv.second;
    }
    for (hipEvent_t e : s->spare_events) (void)hipEventDestroy(e);
    const bool failed = s->failed;
    const std::string msg = s->fail_msg;
    delete s;
    if (failed) return err(DIVER_EHIP, "%s", msg.c_str());
    return DIVER_OK;
}

int diver_server_add_scene(diver_server* s, const diver_grid* grid, const diver_decoder* dec,
                           const diver_render_opts* opts, int64_t scene_bytes,
                           diver_stream_t ready, int* scene_id) {
    if (!s || !grid || !dec || !opts || !scene_id) return err(DIVER_EINVAL, "null argument");
    if (scene_bytes < 0) return err(DIVER_EINVAL, "scene_bytes < 0");
    // an empty launch validates the grid, decoder and options exactly as the frames' launches will
    const int st = diver_render_batch(grid, dec, opts, nullptr, 0, ready);
    if (st != DIVER_OK) return st;
    Scene sc{*grid, dec, *opts, scene_bytes, nullptr};
    hipError_t he = hipEventCreateWithFlags(&sc.ready, hipEventDisableTiming);
    if (he == hipSuccess) he = hipEventRecord(sc.ready, reinterpret_cast<hipStream_t>(ready));
    if (he != hipSuccess) {
        if (sc.ready) (void)hipEventDestroy(sc.ready);
        return err(DIVER_EHIP, "scene event: %s", hipGetErrorString(he));
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->resident + scene_bytes > s->cfg.hbm_budget) {
        (void)hipEventDestroy(sc.ready);
        return err(DIVER_EBUDGET, "the scene needs %lld bytes; %lld of the %lld-byte budget are free",
                   (long long)scene_bytes, (long long)(s->cfg.hbm_budget - s->resident),
                   (long long)s->cfg.hbm_budget);
    }
    s->resident += scene_bytes;
    s->scenes.push_back(sc);
    *scene_id = (int)s->scenes.size() - 1;
    return DIVER_OK;
}

static int submit_frame(diver_server* s, int scene_id, const float* c2w_host, int width,
                        int height, float focal, float fy, float cx, float cy, float* rgba,
                        int32_t* count, diver_stream_t after, int64_t* ticket) {
    if (!s || !c2w_host || !rgba || !ticket) return err(DIVER_EINVAL, "null argument");
    if (width <= 0 || height <= 0 || !(focal > 0) || !(fy >= 0)) return err(DIVER_EINVAL, "bad camera");
    auto* f = new Frame();
    std::memcpy(f->c2w, c2w_host, sizeof(f->c2w));
    f->focal = focal;
    f->fy = fy;
    f->cx = cx;
    f->cy = cy;
    f->rgba = rgba;
    f->count = count;
    f->bytes = (int64_t)width * height * kBytesPerPixel;
    f->ray_ids = (int64_t)((width + 7) / 8) * ((height + 7) / 8) * 64;
    f->after = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->spare_events.empty()) {
            f->after = s->spare_events.back();
            s->spare_events.pop_back();
        }
    }
    hipError_t he = hipSuccess;
    if (!f->after) he = hipEventCreateWithFlags(&f->after, hipEventDisableTiming);
    if (he == hipSuccess) he = hipEventRecord(f->after, reinterpret_cast<hipStream_t>(after));
    if (he != hipSuccess) {
        if (f->after) (void)hipEventDestroy(f->after);
        delete f;
        return err(DIVER_EHIP, "submit event: %s", hipGetErrorString(he));
    }
    std::lock_guard<std::mutex> lk(s->mu);
    int code = DIVER_OK;
    if (scene_id < 0 || scene_id >= (int)s->scenes.size())
        code = err(DIVER_EINVAL, "unknown scene %d", scene_id);
    else if (s->failed)
        code = err(DIVER_ESTATE, "the server has failed: %s", s->fail_msg.c_str());
    else if (s->stopping)
        code = err(DIVER_ESTATE, "the server is closing");
    else if (s->resident + f->bytes > s->cfg.hbm_budget)
        code = err(DIVER_EBUDGET, "a %dx%d frame needs %lld bytes of output; %lld are free", width,
                   height, (long long)f->bytes, (long long)(s->cfg.hbm_budget - s->resident));
    else if (f->ray_ids > s->cfg.max_rays)
        code = err(DIVER_EBUDGET, "%lld ray ids exceed the per-launch limit %lld",
                   (long long)f->ray_ids, (long long)s->cfg.max_rays);
    if (code != DIVER_OK) {
        release_event(s, f->after);
        delete f;
        return code;
    }
    f->large = s->cfg.split_rays > 0 && (int64_t)width * height > s->cfg.split_rays;
    f->info.ticket = s->next_ticket++;
    f->info.scene = scene_id;
    f->info.width = width;
    f->info.height = height;
    f->info.status = -1;
    f->info.launch = -1;
    f->info.t_submit = now_s();
    f->info.t_launch = 0.0;
    f->info.t_done = 0.0;
    s->active[f->info.ticket] = f;
    s->pending.push_back(f);
    *ticket = f->info.ticket;
    s->cv_launch.notify_one();
    return DIVER_OK;
}

int diver_server_submit(diver_server* s, int scene_id, const float* c2w_host, int width,
                        int height, float focal, float* rgba, int32_t* count,
                        diver_stream_t after, int64_t* ticket) {
    return submit_frame(s, scene_id, c2w_host, width, height, focal, 0.0f, 0.0f, 0.0f, rgba, count,
                        after, ticket);
}

int diver_server_submit_pinhole(diver_server* s, int scene_id, const float* c2w_host, int width,
                                int height, float fx, float fy, float cx, float cy, float* rgba,
                                int32_t* count, diver_stream_t after, int64_t* ticket) {
    if (!(fy > 0)) return err(DIVER_EINVAL, "bad camera");
    return submit_frame(s, scene_id, c2w_host, width, height, fx, fy, cx, cy, rgba, count, after,
                        ticket);
}

int diver_server_wait(diver_server* s, int64_t ticket, double timeout_s, diver_frame_info* info) {
    if (!s) return err(DIVER_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return s->finished.count(ticket) || !s->active.count(ticket); };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
        return err(DIVER_ETIMEDOUT, "frame %lld not rendered within %.3f s", (long long)ticket,
                   timeout_s);
    }
    auto it = s->finished.find(ticket);
    if (it == s->finished.end())
        return err(DIVER_EINVAL, "unknown ticket %lld (or already waited for)", (long long)ticket);
    const diver_frame_info fi = it->second;
    s->finished.erase(it);
    if (info) *info = fi;
    if (fi.status != DIVER_OK)
        return err(fi.status, "frame %lld was not rendered: %s", (long long)ticket,
                   s->fail_msg.c_str());
    return DIVER_OK;
}

int64_t diver_server_take_completed(diver_server* s, int64_t* tickets, int64_t max) {
    if (!s || !tickets || max <= 0) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    int64_t n = 0;
    while (n < max && !s->completed.empty()) {
        tickets[n++] = s->completed.front();
        s->completed.pop_front();
    }
    return n;
}

int64_t diver_server_history(diver_server* s, diver_frame_info* out, int64_t max,
                             int64_t* launches) {
    if (!s) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    if (launches) *launches = s->launches;
    if (!out || max <= 0) return 0;
    const int64_t n = std::min<int64_t>(max, (int64_t)s->hist_count);
    const size_t first = (s->hist_next + kHistory - (size_t)n) % kHistory;
    for (int64_t i = 0; i < n; ++i) out[i] = s->history[(first + (size_t)i) % kHistory];
    return n;
}

}  // extern "C"
