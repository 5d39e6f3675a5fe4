// lp_server.hip -- continuous-batching front end for several sensors.
//
// A server owns `slots` contexts (one HIP stream and one preallocated set of
// device buffers each).  Frames are queued as they arrive; lp_server_poll
// collects the batches that finished (non-blocking stream queries), then
// starts queued frames on every free slot as one batch of up to
// max_batch_frames frames, whatever their sensor and image size.  Frames
// that arrive while batches run wait only until a slot frees up, and the
// batches of different slots overlap on the device.  The number of slots is
// what the HBM budget admits.  Every frame of a batch is processed
// independently (per-frame descriptors, CSR-packed inputs), so a frame's
// result is the same alone and in any mix.
//
// No reference counterpart: the reference processor runs one frame per
// point-cloud callback on a single-threaded executor
// (src/processor/src/processor.cpp:878-894); this front end serves that same
// per-frame path to many sensors.
#include <algorithm>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "lp_kernels.h"

namespace {

struct Request {
    lp_segmenter_config cfg;
    lp_cloud cloud;
    lp_frame_outputs out;
    lp_frame_result* result;
    uint64_t tag;
};

struct Slot {
    lp_context* ctx = nullptr;
    std::vector<Request> batch;
    bool busy = false;
};

}  // namespace

struct lp_server {
    lp_server_config cfg{};
    std::vector<Slot> slots;
    std::deque<Request> queue;
    std::deque<uint64_t> done;
    uint64_t batches = 0;
    uint64_t frames = 0;
    uint64_t slot_bytes = 0;
};

namespace {

void finish_failed(lp_server* s, std::vector<Request>& reqs, int status) {
    for (auto& r : reqs) {
        if (r.result) {
            std::memset(r.result, 0, sizeof(*r.result));
            r.result->status = status;
        }
        s->done.push_back(r.tag);
    }
    reqs.clear();
}

// Collects a finished slot: results and outputs of its frames.
int collect(lp_server* s, Slot& sl) {
    const int B = static_cast<int>(sl.batch.size());
    std::vector<lp_frame_outputs> outs(B);
    std::vector<lp_frame_result> res(B);
    for (int k = 0; k < B; ++k) outs[k] = sl.batch[k].out;
    const int st = lp_batch_fetch(sl.ctx, outs.data(), res.data());
    for (int k = 0; k < B; ++k) {
        if (sl.batch[k].result) *sl.batch[k].result = res[k];
        s->done.push_back(sl.batch[k].tag);
    }
    sl.batch.clear();
    sl.busy = false;
    // Per-frame failures are reported in each result; only a device error
    // fails the poll.
    return st == LP_ERR_HIP ? st : LP_OK;
}

int launch(lp_server* s, Slot& sl) {
    const int B = static_cast<int>(sl.batch.size());
    std::vector<lp_segmenter_config> cfgs(B);
    std::vector<lp_cloud> clouds(B);
    for (int k = 0; k < B; ++k) {
        cfgs[k] = sl.batch[k].cfg;
        clouds[k] = sl.batch[k].cloud;
    }
    int st = lp_batch_stage(sl.ctx, cfgs.data(), B, &s->cfg.cluster, B, clouds.data());
    if (st == LP_OK) st = lp_batch_run(sl.ctx, 1);
    if (st != LP_OK) {
        finish_failed(s, sl.batch, st);
        return st == LP_ERR_HIP ? st : LP_OK;
    }
    sl.busy = true;
    s->batches += 1;
    s->frames += static_cast<uint64_t>(B);
    return LP_OK;
}

}  // namespace

extern "C" {

void lp_server_config_default(lp_server_config* cfg) {
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->slots = 3;
    cfg->max_batch_frames = 8;
    lp_capacity_default(&cfg->capacity);
    cfg->capacity.max_image_height = 128;
    cfg->capacity.max_points_per_frame = 262144;
    cfg->hbm_budget_bytes = 0;
    lp_clusterer_config_shipped(&cfg->cluster);
    cfg->noise_filter = 0;
    lp_noise_config_default(&cfg->noise);
    cfg->oriented_boxes = 0;
    cfg->label_clouds = 0;
    cfg->vehicle_boxes = 0;
    cfg->noise_search = LP_NOISE_RULE;
}

int lp_server_create(const lp_server_config* cfg, lp_server** out) {
    if (!cfg || !out) return LP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (cfg->slots < 1 || cfg->max_batch_frames < 1) return LP_ERR_INVALID_ARGUMENT;
    auto* s = new lp_server();
    s->cfg = *cfg;
    lp_capacity cap = cfg->capacity;
    cap.max_frames = cfg->max_batch_frames;
    cap.streams = 1;  // slots already run side by side
    for (int k = 0; k < cfg->slots; ++k) {
        lp_context* c = nullptr;
        const int st = lp_context_create(&cap, &c);
        if (st != LP_OK) {
            lp_server_destroy(s);
            return st;
        }
        // Admission by HBM budget: keep a slot only while all slots' buffers fit.
        if (cfg->hbm_budget_bytes && (s->slots.size() + 1) * c->arena_bytes > cfg->hbm_budget_bytes) {
            lp_context_destroy(c);
            break;
        }
        s->slot_bytes = c->arena_bytes;
        int ns = lp_set_noise_search(c, cfg->noise_search);
        if (ns == LP_OK && cfg->noise_filter) ns = lp_set_noise_filter(c, &cfg->noise);
        if (ns != LP_OK) {
            lp_context_destroy(c);
            lp_server_destroy(s);
            return ns;
        }
        lp_set_oriented_boxes(c, cfg->oriented_boxes);
        lp_set_vehicle_boxes(c, cfg->vehicle_boxes);
        const int cl = lp_set_label_clouds(c, cfg->label_clouds);
        if (cl != LP_OK) {
            lp_context_destroy(c);
            lp_server_destroy(s);
            return cl;
        }
        Slot sl;
        sl.ctx = c;
        s->slots.push_back(std::move(sl));
    }
    if (s->slots.empty()) {
        lp_server_destroy(s);
        return LP_ERR_CAPACITY;
    }
    *out = s;
    return LP_OK;
}

int lp_server_destroy(lp_server* s) {
    if (!s) return LP_OK;
    for (auto& sl : s->slots)
        if (sl.ctx) lp_context_destroy(sl.ctx);
    delete s;
    return LP_OK;
}

int lp_server_submit(lp_server* s, const lp_segmenter_config* cfg, const lp_cloud* cloud,
                     const lp_frame_outputs* outputs, lp_frame_result* result, uint64_t tag) {
    if (!s || !cfg || !cloud || !outputs) return LP_ERR_INVALID_ARGUMENT;
    // A frame that staging would refuse is refused here, so it never fails
    // the other frames of the batch it would have joined.  Every slot has
    // the same capacity.
    const int st = lp::validate_frame(s->slots.front().ctx, *cfg, *cloud);
    if (st != LP_OK) return st;
    s->queue.push_back(Request{*cfg, *cloud, *outputs, result, tag});
    return LP_OK;
}

int lp_server_poll(lp_server* s, uint64_t* done, int32_t max_done) {
    if (!s || (max_done > 0 && !done)) return LP_ERR_INVALID_ARGUMENT;
    int err = LP_OK;
    for (auto& sl : s->slots) {
        if (!sl.busy) continue;
        const int q = lp_query(sl.ctx);
        if (q < 0) return q;
        if (q == 1) {
            const int st = collect(s, sl);
            if (st != LP_OK) err = st;
        }
    }
    for (auto& sl : s->slots) {
        if (s->queue.empty()) break;
        if (sl.busy) continue;
        const int B = std::min<int>(s->cfg.max_batch_frames, static_cast<int>(s->queue.size()));
        for (int k = 0; k < B; ++k) {
            sl.batch.push_back(s->queue.front());
            s->queue.pop_front();
        }
        const int st = launch(s, sl);
        if (st != LP_OK) err = st;
    }
    if (err != LP_OK) return err;
    int n = 0;
    while (n < max_done && !s->done.empty()) {
        done[n++] = s->done.front();
        s->done.pop_front();
    }
    return n;
}

int lp_server_stats_get(lp_server* s, lp_server_stats* st) {
    if (!s || !st) return LP_ERR_INVALID_ARGUMENT;
    std::memset(st, 0, sizeof(*st));
    st->slots = static_cast<int32_t>(s->slots.size());
    for (const auto& sl : s->slots)
        if (sl.busy) {
            st->busy_slots += 1;
            st->in_flight += static_cast<int32_t>(sl.batch.size());
        }
    st->queued = static_cast<int32_t>(s->queue.size());
    st->slot_bytes = s->slot_bytes;
    st->batches = s->batches;
    st->frames = s->frames;
    return LP_OK;
}

}  // extern "C"
