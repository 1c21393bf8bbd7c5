// server.cpp -- the serving loop of the stitch path: streams of different rigs
// admitted against an HBM budget, their frames batched continuously across
// streams (one launch per pyramid stage per batch, pano_batch_render), with
// submit-to-ready latency per frame.
//
// This replaces the per-stream worker loop of the reference's live task
// (PanoramaLiveStreamTask2.cpp:340-350, 979-983: one render object and one
// thread per stream, each calling CudaPanoramaRender2::render on its own).
// Here one launcher thread keeps up to max_inflight batches queued on the
// server's HIP stream: whenever a batch may be launched, it takes the oldest
// pending frame of every stream that has one (one frame per stream and batch,
// since a render's planes serve one frame at a time; consecutive batches of
// one stream are ordered by the HIP stream) and launches them together; a
// completion thread waits on each batch's event in order and timestamps its
// frames.  Frames that arrive while batches run join the next batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/panohip.h"
#include "pano_dev.hpp"

using pano::kMaxBatch;
using pano::kMaxSrc;

namespace {

using Clock = std::chrono::steady_clock;

struct Frame {
    int id = 0;
    uint64_t tag = 0;
    std::vector<const uint8_t*> srcs;
    std::vector<size_t> steps;
    std::vector<uint8_t> luts;   // empty: no correction
    uint8_t* dst = nullptr;
    size_t dst_step = 0;
    Clock::time_point submitted;
};

struct Stream {
    pano_render* r = nullptr;
    int n = 0;
    int64_t bytes = 0;
    bool active = false;
    std::deque<Frame> pending;
    int inflight = 0;
};

struct Batch {
    hipEvent_t ev = nullptr;
    std::vector<Frame> frames;
    int status = PANO_OK;
};

}  // namespace

struct pano_server {
    int64_t budget = 0, used = 0;
    int max_inflight = 2;
    hipStream_t stream = nullptr;
    pano_batch* batch = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    std::vector<Stream> streams;
    std::deque<Batch> inflight;
    std::vector<hipEvent_t> spare_events;
    std::vector<pano_server_done> done;
    std::thread launcher, completer;
    bool stop = false;
    int64_t frames_done = 0, batches = 0;

    bool can_launch() const
    {
        if ((int)inflight.size() >= max_inflight) return false;
        for (const Stream& s : streams)
            if (s.active && !s.pending.empty()) return true;
        return false;
    }

    void launch_loop()
    {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv.wait(lk, [&] { return stop || can_launch(); });
            if (stop) return;
            // the oldest pending frame of every stream
            std::vector<int> pick;
            for (int id = 0; id < (int)streams.size(); id++) {
                const Stream& s = streams[id];
                if (s.active && !s.pending.empty()) pick.push_back(id);
            }
            std::sort(pick.begin(), pick.end(), [&](int a, int b) {
                return streams[a].pending.front().submitted < streams[b].pending.front().submitted;
            });
            Batch bt;
            std::vector<pano_render*> rs;
            int cams = 0, shape = -1;
            for (int id : pick) {
                Stream& s = streams[id];
                if ((int)rs.size() == kMaxBatch || cams + s.n > kMaxSrc) break;
                // one launch per stage needs one down tile shape: others wait for a later batch
                const int sh = pano_render_tile_shape(s.r);
                if (shape < 0) shape = sh;
                else if (sh != shape) continue;
                rs.push_back(s.r);
                cams += s.n;
                bt.frames.push_back(std::move(s.pending.front()));
                s.pending.pop_front();
                s.inflight++;
            }
            if (!spare_events.empty()) {
                bt.ev = spare_events.back();
                spare_events.pop_back();
            }
            lk.unlock();
            std::vector<const uint8_t*> srcs;
            std::vector<size_t> steps;
            std::vector<const uint8_t*> luts;
            std::vector<uint8_t*> dsts;
            std::vector<size_t> dsteps;
            bool any_lut = false;
            for (const Frame& f : bt.frames) {
                srcs.insert(srcs.end(), f.srcs.begin(), f.srcs.end());
                steps.insert(steps.end(), f.steps.begin(), f.steps.end());
                luts.push_back(f.luts.empty() ? nullptr : f.luts.data());
                any_lut |= !f.luts.empty();
                dsts.push_back(f.dst);
                dsteps.push_back(f.dst_step);
            }
            int st = PANO_OK;
            if (!bt.ev && hipEventCreateWithFlags(&bt.ev, hipEventDisableTiming) != hipSuccess)
                st = PANO_ERR_HIP;
            if (!st) st = pano_batch_set(batch, (int)rs.size(), rs.data());
            if (!st)
                st = pano_batch_render(batch, srcs.data(), steps.data(), any_lut ? luts.data() : nullptr,
                                       dsts.data(), dsteps.data(), stream);
            if (!st && hipEventRecord(bt.ev, stream) != hipSuccess) st = PANO_ERR_HIP;
            bt.status = st;
            lk.lock();
            batches++;
            inflight.push_back(std::move(bt));
            cv.notify_all();
        }
    }

    void complete_loop()
    {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv.wait(lk, [&] { return !inflight.empty() || (stop && inflight.empty()); });
            if (inflight.empty()) return;   // stopping with nothing in flight
            hipEvent_t ev = inflight.front().ev;
            const bool failed = inflight.front().status != PANO_OK;
            lk.unlock();
            int st = PANO_OK;
            if (!failed && hipEventSynchronize(ev) != hipSuccess) st = PANO_ERR_HIP;
            const Clock::time_point t = Clock::now();
            lk.lock();
            Batch bt = std::move(inflight.front());
            inflight.pop_front();
            if (bt.status == PANO_OK) bt.status = st;
            for (const Frame& f : bt.frames) {
                pano_server_done d{};
                d.stream_id = f.id;
                d.tag = f.tag;
                d.status = bt.status;
                d.latency_ms = std::chrono::duration<double, std::milli>(t - f.submitted).count();
                d.batch_frames = (int)bt.frames.size();
                done.push_back(d);
                streams[f.id].inflight--;
                frames_done++;
            }
            if (bt.ev) spare_events.push_back(bt.ev);
            cv.notify_all();
        }
    }
};

extern "C" {

int pano_server_create(pano_server** out, int64_t hbm_budget_bytes, int max_inflight)
{
    if (!out || hbm_budget_bytes <= 0 || max_inflight < 1 || max_inflight > 8) return PANO_ERR_ARG;
    pano_server* s = new (std::nothrow) pano_server();
    if (!s) return PANO_ERR_NOMEM;
    s->budget = hbm_budget_bytes;
    s->max_inflight = max_inflight;
    int e = pano_batch_create(&s->batch);
    if (!e && hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) e = PANO_ERR_HIP;
    if (e) {
        if (s->batch) pano_batch_destroy(s->batch);
        delete s;
        return e;
    }
    s->launcher = std::thread([s] { s->launch_loop(); });
    s->completer = std::thread([s] { s->complete_loop(); });
    *out = s;
    return PANO_OK;
}

void pano_server_destroy(pano_server* s)
{
    if (!s) return;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    s->launcher.join();
    s->completer.join();   // drains the batches in flight
    (void)hipStreamSynchronize(s->stream);
    for (hipEvent_t e : s->spare_events) (void)hipEventDestroy(e);
    pano_batch_destroy(s->batch);
    (void)hipStreamDestroy(s->stream);
    delete s;
}

int pano_server_admit(pano_server* s, pano_render* r, int* stream_id)
{
    if (!s || !r || !stream_id) return PANO_ERR_ARG;
    if (!pano_render_fused(r)) return PANO_ERR_UNSUPPORTED;
    const int64_t bytes = pano_render_mem_bytes(r);
    std::unique_lock<std::mutex> lk(s->mu);
    for (const Stream& t : s->streams)
        if (t.active && t.r == r) return PANO_ERR_STATE;   // one stream per render
    if (s->used + bytes > s->budget) return PANO_ERR_BUDGET;
    int id = -1;
    for (int k = 0; k < (int)s->streams.size(); k++)
        if (!s->streams[k].active && s->streams[k].inflight == 0 && s->streams[k].pending.empty()) {
            id = k;
            break;
        }
    if (id < 0) {
        id = (int)s->streams.size();
        s->streams.emplace_back();
    }
    Stream& t = s->streams[id];
    t.r = r;
    t.n = pano_render_num_images(r);
    t.bytes = bytes;
    t.active = true;
    s->used += bytes;
    *stream_id = id;
    return PANO_OK;
}

int pano_server_retire(pano_server* s, int stream_id)
{
    if (!s) return PANO_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (stream_id < 0 || stream_id >= (int)s->streams.size() || !s->streams[stream_id].active)
        return PANO_ERR_ARG;
    // its submitted frames still run; the render is released once they are done
    s->cv.wait(lk, [&] {
        const Stream& t = s->streams[stream_id];
        return t.pending.empty() && t.inflight == 0;
    });
    Stream& t = s->streams[stream_id];
    t.active = false;
    t.r = nullptr;
    s->used -= t.bytes;
    t.bytes = 0;
    return PANO_OK;
}

int pano_server_submit(pano_server* s, int stream_id, const uint8_t* const* srcs,
                       const size_t* src_steps, const uint8_t* luts, uint8_t* dst, size_t dst_step,
                       uint64_t tag)
{
    if (!s || !srcs || !src_steps || !dst) return PANO_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (stream_id < 0 || stream_id >= (int)s->streams.size() || !s->streams[stream_id].active)
        return PANO_ERR_ARG;
    Stream& t = s->streams[stream_id];
    Frame f;
    f.id = stream_id;
    f.tag = tag;
    f.srcs.assign(srcs, srcs + t.n);
    f.steps.assign(src_steps, src_steps + t.n);
    for (int c = 0; c < t.n; c++)
        if (!srcs[c]) return PANO_ERR_ARG;
    if (luts) f.luts.assign(luts, luts + (size_t)t.n * 768);
    f.dst = dst;
    f.dst_step = dst_step;
    f.submitted = Clock::now();
    t.pending.push_back(std::move(f));
    s->cv.notify_all();
    return PANO_OK;
}

int pano_server_poll(pano_server* s, pano_server_done* out, int max_out, int wait_ms)
{
    if (!s || (!out && max_out > 0) || max_out < 0) return PANO_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (wait_ms > 0 && s->done.empty())
        s->cv.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
    const int k = std::min(max_out, (int)s->done.size());
    for (int i = 0; i < k; i++) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    return k;
}

int pano_server_drain(pano_server* s)
{
    if (!s) return PANO_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    // a frame taken into a batch counts in its stream's inflight from then until
    // the completion thread has reported it, including while it is being launched
    s->cv.wait(lk, [&] {
        for (const Stream& t : s->streams)
            if (!t.pending.empty() || t.inflight) return false;
        return true;
    });
    return PANO_OK;
}

int pano_server_stats(pano_server* s, int64_t* used_bytes, int64_t* frames_done, int64_t* batches)
{
    if (!s) return PANO_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (used_bytes) *used_bytes = s->used;
    if (frames_done) *frames_done = s->frames_done;
    if (batches) *batches = s->batches;
    return PANO_OK;
}

}  // extern "C"
