// Segment server: many antenna streams, per-stream state resident in HBM,
// admission by a device input arena, continuous batching over two HIP
// streams with one persistent (graph-replayed) dedispersion launch per
// iteration.  Host code only: every device operation goes through the
// library's own C ABI.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"

namespace {

constexpr int64_t kAlign = 256;
int64_t align_up(int64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over [0, size) with a coalescing free list.
class Arena {
  public:
    void reset(int64_t size) {
        free_.clear();
        if (size > 0) free_[0] = size;
        used_ = peak_ = 0;
    }
    bool alloc(int64_t n, int64_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const int64_t rest = it->second - n, start = it->first + n;
            free_.erase(it);
            if (rest > 0) free_[start] = rest;
            used_ += n;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(int64_t off, int64_t n) {
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t peak() const { return peak_; }

  private:
    std::map<int64_t, int64_t> free_;
    int64_t used_ = 0, peak_ = 0;
};

// Per-stream results block, same layout on the device and in pinned host
// memory: ave, sigma, peak [ndm] f32; peak_t [ndm] i32; cand_dm, cand_t
// [max_cand] i32; cand_snr [max_cand] f32; count i32.
struct ResLayout {
    size_t ave, sigma, peak, peak_t, cand_dm, cand_t, cand_snr, count, bytes;
    void set(int ndm, int max_cand) {
        size_t o = 0;
        auto take = [&](size_t n) {
            const size_t r = o;
            o += (n * 4 + 255) / 256 * 256;
            return r;
        };
        ave = take(ndm);
        sigma = take(ndm);
        peak = take(ndm);
        peak_t = take(ndm);
        cand_dm = take(max_cand);
        cand_t = take(max_cand);
        cand_snr = take(max_cand);
        count = take(1);
        bytes = o;
    }
};

struct Request {
    int64_t tag = 0;
    int stream = 0;
    const uint8_t* pol[2] = {nullptr, nullptr};
    int64_t nbytes = 0;
    bool device = false;  // payload already in device memory
    vfrb_server_output out{};
    double arrival = 0.0;
};

}  // namespace

struct SrvStream {
    int id = 0, ndm = 0, max_delay = 0;
    vfrb_search_plan* plan = nullptr;
    float* ring = nullptr;  // raw filterbank ring, f32 [ring_rows][nchans]
    int64_t ring_rows = 0;
    uint8_t* carry = nullptr;  // [2][row_bytes] payload bytes short of a row
    int64_t carry_bytes = 0, total_rows = 0;
    float* out = nullptr;  // f32 [ndm][ld_out]
    int64_t ld_out = 0;
    ResLayout lay{};
    unsigned char* d_res = nullptr;
    unsigned char* h_res = nullptr;  // pinned
    bool busy = false;
    uint64_t key = 0;  // filterbank shape + delay parameters + DM list (saved states carry it)
};

namespace {

struct Item {
    Request req;
    SrvStream* st = nullptr;
    bool direct = false;  // device input read in place (no arena region)
    int64_t off[2] = {0, 0}, region = 0;
    int64_t rows = 0, sample0 = 0;
    int win_rows = 0, nt = 0;
    vfrb_status status = VFRB_OK;
};

struct Lane {
    hipStream_t s = nullptr;
    // One batch per dedispersion tile, created on first use with that tile
    // fixed.  A segment's jobs go to the batch of its plan's own tile, so
    // its arithmetic (tile, summation order) is the same in any mix.
    std::vector<vfrb_batch*> batches;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_copy = nullptr;  // this lane's host-to-device copies done (copy stream)
    bool busy = false;
    bool issued = false;  // ev1 recorded for the current iteration
    double start = 0.0;
    std::vector<Item> items;
};

}  // namespace

struct vfrb_server {
    vfrb_server_desc d{};
    int device = 0;
    int64_t row_bytes = 0, max_seg_rows = 0;
    vfrb_fb_plan* fbp = nullptr;
    unsigned char* arena_base = nullptr;
    Arena arena;
    std::vector<std::unique_ptr<SrvStream>> streams;
    std::vector<Lane> lanes;
    // With two lanes, every host-to-device copy goes through this one stream
    // in issue order: concurrent copies from both lanes would share the PCIe
    // link and delay each other, while one queue keeps it busy back to back
    // and the lanes' kernels overlap the next copies.
    hipStream_t copy_s = nullptr;
    std::deque<Request> pending;
    std::deque<vfrb_server_result> completed;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    bool stop = false, mem_blocked = false;
    int64_t submitted = 0, finished = 0, iterations = 0, batched_segments = 0;
    double busy_s = 0.0;
    std::chrono::steady_clock::time_point t0;
    std::thread sched, compl_;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

void free_stream(SrvStream* st) {
    if (st->plan) vfrb_search_plan_destroy(st->plan);
    (void)hipFree(st->ring);
    (void)hipFree(st->carry);
    (void)hipFree(st->out);
    (void)hipFree(st->d_res);
    (void)hipHostFree(st->h_res);
}

// Device input is read in place when the filterbank's 4-byte-aligned reads
// of the rows after the carried partial row start aligned in both pols.
bool direct_ok(const Request& r, int64_t carry, int64_t rb) {
    const int64_t head = carry > 0 ? rb - carry : 0;
    if (head > r.nbytes) return false;  // the segment does not complete the row: copy path
    return ((uintptr_t)(r.pol[0] + head) % 4) == 0 && ((uintptr_t)(r.pol[1] + head) % 4) == 0;
}

bool any_lane_free(const vfrb_server* S) {
    for (const Lane& l : S->lanes)
        if (!l.busy) return true;
    return false;
}

bool any_ready(const vfrb_server* S) {
    for (const Request& r : S->pending)
        if (!S->streams[r.stream]->busy) return true;
    return false;
}

// Enqueues one iteration on lane L.  Called by the scheduler thread without
// the lock: the segments' streams are marked busy, so nothing else touches
// their state until the iteration completes.
void issue(vfrb_server* S, Lane& L) {
    hipStream_t s = L.s;
    (void)hipEventRecord(L.ev0, s);
    const int64_t rb = S->row_bytes;
    // host input: the copies first, on the shared copy stream; the lane waits
    bool host_copied = false;
    if (S->copy_s) {
        for (Item& it : L.items) {
            if (it.req.device) continue;
            const int64_t C = it.st->carry_bytes;
            for (int p = 0; p < 2 && it.status == VFRB_OK; p++)
                it.status = vfrb_memcpy_h2d(S->arena_base + it.off[p] + C, it.req.pol[p], (size_t)it.req.nbytes,
                                            S->copy_s);
            host_copied = true;
        }
        if (host_copied) {
            hipError_t he = hipEventRecord(L.ev_copy, S->copy_s);
            if (he == hipSuccess) he = hipStreamWaitEvent(s, L.ev_copy, 0);
            if (he != hipSuccess) {
                const vfrb_status e = vfrb::hip_fail(he, "server copy queue", __FILE__, __LINE__);
                for (Item& it : L.items)
                    if (it.status == VFRB_OK) it.status = e;
            }
        }
    }
    for (Item& it : L.items) {
        if (it.status != VFRB_OK) continue;
        SrvStream* st = it.st;
        const int64_t C = st->carry_bytes, tot = C + it.req.nbytes;
        const int64_t R = tot / rb, newc = tot - R * rb;
        vfrb_status e = VFRB_OK;
        if (it.direct) {
            // Device input read in place: the carried partial row is
            // completed in the carry buffer and filterbanked from there, the
            // rest straight from the caller's payload; the same rows through
            // the same kernel as the arena copy below.
            const int64_t head = C > 0 ? std::min(rb - C, it.req.nbytes) : 0;
            for (int p = 0; p < 2 && e == VFRB_OK && head > 0; p++)
                e = vfrb_memcpy_d2d(st->carry + p * rb + C, it.req.pol[p], (size_t)head, s);
            int64_t done = 0;
            if (e == VFRB_OK && C > 0 && C + head == rb) {
                e = vfrb_fb_execute(S->fbp, st->carry, st->carry + rb, 1, st->ring, st->ring_rows, st->total_rows,
                                    s);
                done = 1;
            }
            if (e == VFRB_OK && R > done)
                e = vfrb_fb_execute(S->fbp, it.req.pol[0] + head, it.req.pol[1] + head, R - done, st->ring,
                                    st->ring_rows, st->total_rows + done, s);
            // the new partial row: the segment's last newc bytes once a row
            // completed; otherwise (only from an empty carry) all its bytes
            for (int p = 0; p < 2 && e == VFRB_OK; p++) {
                if (R > 0 && newc > 0)
                    e = vfrb_memcpy_d2d(st->carry + p * rb, it.req.pol[p] + (it.req.nbytes - newc), (size_t)newc,
                                        s);
                else if (R == 0 && head == 0 && it.req.nbytes > 0)
                    e = vfrb_memcpy_d2d(st->carry + p * rb + C, it.req.pol[p], (size_t)it.req.nbytes, s);
            }
        } else {
            unsigned char* reg[2] = {S->arena_base + it.off[0], S->arena_base + it.off[1]};
            for (int p = 0; p < 2 && e == VFRB_OK; p++) {
                if (C > 0) e = vfrb_memcpy_d2d(reg[p], st->carry + p * rb, (size_t)C, s);
                if (e == VFRB_OK && it.req.device)
                    e = vfrb_memcpy_d2d(reg[p] + C, it.req.pol[p], (size_t)it.req.nbytes, s);
                else if (e == VFRB_OK && !host_copied)
                    e = vfrb_memcpy_h2d(reg[p] + C, it.req.pol[p], (size_t)it.req.nbytes, s);
            }
            if (e == VFRB_OK && R > 0)
                e = vfrb_fb_execute(S->fbp, reg[0], reg[1], R, st->ring, st->ring_rows, st->total_rows, s);
            for (int p = 0; p < 2 && e == VFRB_OK; p++)
                if (newc > 0) e = vfrb_memcpy_d2d(st->carry + p * rb, reg[p] + R * rb, (size_t)newc, s);
        }
        if (e != VFRB_OK) {
            it.status = e;
            continue;
        }
        st->carry_bytes = newc;
        st->total_rows += R;
        it.rows = R;
        const int64_t W = std::min<int64_t>(st->total_rows, st->max_delay + R);
        it.nt = (int)std::max<int64_t>(0, W - st->max_delay);
        if (it.nt < 2) {  // calcstats needs two samples
            it.nt = 0;
            continue;
        }
        it.win_rows = (int)W;
        it.sample0 = st->total_rows - W;
        it.status = vfrb_search_load(st->plan, st->ring, st->ring_rows, it.sample0, (int32_t)W, 1, s);
        if (it.status != VFRB_OK) it.nt = 0;
    }
    // dedispersion: one persistent launch for the whole iteration, or one
    // launch per segment on the plain path
    std::vector<vfrb_search_plan*> plans;
    std::vector<int32_t> nrows;
    std::vector<float*> outs;
    std::vector<int64_t> lds;
    std::vector<Item*> jobs;
    for (Item& it : L.items) {
        if (it.status != VFRB_OK || it.nt == 0) continue;
        jobs.push_back(&it);
        plans.push_back(it.st->plan);
        nrows.push_back(it.win_rows);
        outs.push_back(it.st->out);
        lds.push_back(it.st->ld_out);
    }
    if (!jobs.empty()) {
        if (S->d.batched) {
            // one persistent launch per tile in use (one in the usual case:
            // streams whose DM lists all fit the default tile)
            std::vector<int32_t> tile(jobs.size());
            for (size_t i = 0; i < jobs.size(); i++) tile[i] = vfrb_search_get_tile(plans[i]);
            std::vector<bool> done(jobs.size(), false);
            for (size_t i = 0; i < jobs.size(); i++) {
                if (done[i]) continue;
                const int32_t t = tile[i];
                std::vector<size_t> grp;
                for (size_t k = i; k < jobs.size(); k++)
                    if (!done[k] && tile[k] == t) {
                        grp.push_back(k);
                        done[k] = true;
                    }
                vfrb_status e = VFRB_OK;
                if (t < 0 || t >= (int32_t)L.batches.size()) e = VFRB_ERR_INVALID;
                if (e == VFRB_OK && !L.batches[t])
                    e = vfrb_batch_create(&L.batches[t], S->d.storage, t, S->d.max_streams, 1 << 22);
                std::vector<vfrb_search_plan*> gp;
                std::vector<int32_t> gr;
                std::vector<float*> go;
                std::vector<int64_t> gl;
                for (size_t k : grp) {
                    gp.push_back(plans[k]);
                    gr.push_back(nrows[k]);
                    go.push_back(outs[k]);
                    gl.push_back(lds[k]);
                }
                if (e == VFRB_OK)
                    e = vfrb_batch_set(L.batches[t], (int32_t)grp.size(), gp.data(), gr.data(), go.data(),
                                       gl.data(), nullptr, s);
                if (e == VFRB_OK) e = vfrb_batch_graph_launch(L.batches[t], s);
                if (e != VFRB_OK)
                    for (size_t k : grp) jobs[k]->status = e;
            }
        } else {
            for (Item* it : jobs)
                it->status = vfrb_search_dedisperse(it->st->plan, it->win_rows, it->st->out, it->st->ld_out, s);
        }
    }
    for (Item* it : jobs) {
        if (it->status != VFRB_OK) continue;
        SrvStream* st = it->st;
        const ResLayout& y = st->lay;
        unsigned char* r = st->d_res;
        vfrb_status e = vfrb_dm_stats(st->out, st->ld_out, st->ndm, it->nt, S->d.acc, nullptr,
                                      (float*)(r + y.ave), (float*)(r + y.sigma), (float*)(r + y.peak),
                                      (int32_t*)(r + y.peak_t), s);
        if (e == VFRB_OK)
            e = vfrb_candidates(st->out, st->ld_out, st->ndm, it->nt, (float*)(r + y.ave),
                                (float*)(r + y.sigma), S->d.threshold, (int32_t*)(r + y.cand_dm),
                                (int32_t*)(r + y.cand_t), (float*)(r + y.cand_snr), S->d.max_cand,
                                (int32_t*)(r + y.count), s);
        if (e == VFRB_OK) e = vfrb_memcpy_d2h(st->h_res, st->d_res, y.bytes, s);
        it->status = e;
    }
    (void)hipEventRecord(L.ev1, s);
}

// Builds the result of one finished segment and copies its outputs into the
// caller's buffers.
vfrb_server_result finish_item(vfrb_server* S, Item& it, const Lane& L, double done) {
    vfrb_server_result r{};
    r.tag = it.req.tag;
    r.stream_id = it.req.stream;
    r.status = it.status;
    r.rows = it.rows;
    r.sample0 = it.sample0;
    r.nt = it.status == VFRB_OK ? it.nt : 0;
    r.best_dm = -1;
    r.best_t = -1;
    r.batch_size = (int32_t)L.items.size();
    r.arrival_s = it.req.arrival;
    r.start_s = L.start;
    r.done_s = done;
    if (r.nt == 0) return r;
    SrvStream* st = it.st;
    const ResLayout& y = st->lay;
    const unsigned char* h = st->h_res;
    const float* peak = (const float*)(h + y.peak);
    const int32_t* peak_t = (const int32_t*)(h + y.peak_t);
    for (int d = 0; d < st->ndm; d++)
        if (r.best_dm < 0 || peak[d] > r.best_snr) {
            r.best_dm = d;
            r.best_snr = peak[d];
            r.best_t = (int32_t)(peak_t[d] + it.sample0);
        }
    const vfrb_server_output& o = it.req.out;
    if (o.peak) std::memcpy(o.peak, peak, sizeof(float) * st->ndm);
    if (o.ave) std::memcpy(o.ave, h + y.ave, sizeof(float) * st->ndm);
    if (o.sigma) std::memcpy(o.sigma, h + y.sigma, sizeof(float) * st->ndm);
    if (o.peak_t)
        for (int d = 0; d < st->ndm; d++) o.peak_t[d] = (int32_t)(peak_t[d] + it.sample0);
    r.ncand = *(const int32_t*)(h + y.count);
    const int n = std::min(r.ncand, S->d.max_cand);
    const int32_t* cd = (const int32_t*)(h + y.cand_dm);
    const int32_t* ct = (const int32_t*)(h + y.cand_t);
    const float* cs = (const float*)(h + y.cand_snr);
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) {
        return cd[a] != cd[b] ? cd[a] < cd[b] : ct[a] < ct[b];
    });
    for (int i = 0; i < n; i++) {
        if (o.cand_dm) o.cand_dm[i] = cd[idx[i]];
        if (o.cand_t) o.cand_t[i] = (int32_t)(ct[idx[i]] + it.sample0);
        if (o.cand_snr) o.cand_snr[i] = cs[idx[i]];
    }
    return r;
}

void scheduler_loop(vfrb_server* S) {
    (void)hipSetDevice(S->device);
    const int max_batch = S->d.batched ? std::max(1, S->d.max_batch) : 1;
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        S->cv_sched.wait(lk, [&] {
            return S->stop || (any_lane_free(S) && !S->mem_blocked && any_ready(S));
        });
        if (S->stop) return;
        Lane* L = nullptr;
        for (Lane& l : S->lanes)
            if (!l.busy) {
                L = &l;
                break;
            }
        L->items.clear();
        for (auto it = S->pending.begin(); it != S->pending.end() && (int)L->items.size() < max_batch;) {
            SrvStream* st = S->streams[it->stream].get();
            if (st->busy) {
                ++it;
                continue;
            }
            Item item;
            item.direct = it->device && direct_ok(*it, st->carry_bytes, S->row_bytes);
            if (!item.direct) {
                const int64_t reg = align_up(st->carry_bytes + it->nbytes);
                if (!S->arena.alloc(reg, &item.off[0])) break;  // arrival order: no overtaking
                if (!S->arena.alloc(reg, &item.off[1])) {
                    S->arena.release(item.off[0], reg);
                    break;
                }
                item.region = reg;
            }
            item.req = *it;
            item.st = st;
            st->busy = true;
            L->items.push_back(item);
            it = S->pending.erase(it);
        }
        if (L->items.empty()) {
            S->mem_blocked = true;  // cleared when an iteration frees arena space
            continue;
        }
        L->busy = true;
        L->issued = false;
        L->start = S->now();
        S->iterations++;
        S->batched_segments += (int64_t)L->items.size();
        lk.unlock();
        issue(S, *L);
        lk.lock();
        L->issued = true;
        S->cv_done.notify_all();  // the completer watches this lane now
    }
}

void completer_loop(vfrb_server* S) {
    (void)hipSetDevice(S->device);
    for (;;) {
        std::vector<Lane*> busy;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_done.wait(lk, [&] {
                if (S->stop) return true;
                for (Lane& l : S->lanes)
                    if (l.busy && l.issued) return true;
                return false;
            });
            bool any = false;
            for (Lane& l : S->lanes) any = any || l.busy;
            if (S->stop && !any) return;
            for (Lane& l : S->lanes)
                if (l.busy && l.issued) busy.push_back(&l);
        }
        bool progressed = false;
        for (Lane* L : busy) {
            const hipError_t q = hipEventQuery(L->ev1);
            if (q == hipErrorNotReady) continue;
            const double done = S->now();
            float ms = 0.0f;
            if (q == hipSuccess) (void)hipEventElapsedTime(&ms, L->ev0, L->ev1);
            std::vector<vfrb_server_result> res;
            for (Item& it : L->items) {
                if (q != hipSuccess && it.status == VFRB_OK) it.status = VFRB_ERR_HIP;
                res.push_back(finish_item(S, it, *L, done));
            }
            std::lock_guard<std::mutex> lk(S->mu);
            for (Item& it : L->items) {
                if (it.region > 0) {
                    S->arena.release(it.off[0], it.region);
                    S->arena.release(it.off[1], it.region);
                }
                it.st->busy = false;
            }
            for (auto& r : res) S->completed.push_back(r);
            S->finished += (int64_t)res.size();
            S->busy_s += ms * 1e-3;
            L->items.clear();
            L->busy = false;
            L->issued = false;
            L->start = 0.0;
            S->mem_blocked = false;
            progressed = true;
            S->cv_sched.notify_all();
            S->cv_done.notify_all();
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

void server_free(vfrb_server* S) {
    for (auto& st : S->streams) free_stream(st.get());
    for (Lane& l : S->lanes) {
        for (vfrb_batch* b : l.batches)
            if (b) vfrb_batch_destroy(b);
        if (l.ev0) (void)hipEventDestroy(l.ev0);
        if (l.ev1) (void)hipEventDestroy(l.ev1);
        if (l.ev_copy) (void)hipEventDestroy(l.ev_copy);
        if (l.s) (void)hipStreamDestroy(l.s);
    }
    if (S->copy_s) (void)hipStreamDestroy(S->copy_s);
    if (S->fbp) vfrb_fb_plan_destroy(S->fbp);
    (void)hipFree(S->arena_base);
    delete S;
}

}  // namespace

extern "C" {

vfrb_status vfrb_server_create(vfrb_server** server, const vfrb_server_desc* d) {
    VFRB_REQUIRE(server && d, VFRB_ERR_INVALID, "NULL argument");
    *server = nullptr;
    VFRB_REQUIRE(d->twonchan > 0 && d->nave > 0 && d->nchans > 0 && d->nskip >= 0, VFRB_ERR_INVALID,
                 "bad filterbank shape");
    VFRB_REQUIRE(d->tsamp > 0.0 && d->chanbw > 0.0 && d->fch1 > 0.0, VFRB_ERR_INVALID, "bad band");
    VFRB_REQUIRE(d->max_streams > 0 && d->max_segment_bytes > 0 && d->arena_bytes > 0 &&
                     d->max_cand >= 0, VFRB_ERR_INVALID, "bad capacities");
    VFRB_REQUIRE(d->acc == VFRB_ACC_F64 || d->acc == VFRB_ACC_F32C, VFRB_ERR_INVALID, "acc=%d", d->acc);
    VFRB_REQUIRE(d->tile_config >= -1 && d->tile_config < vfrb_search_tile_count(), VFRB_ERR_INVALID,
                 "tile_config=%d", d->tile_config);
    vfrb_server* S = new vfrb_server();
    S->d = *d;
    S->t0 = std::chrono::steady_clock::now();
    S->row_bytes = (int64_t)d->twonchan * d->nave;
    S->max_seg_rows = d->max_segment_bytes / S->row_bytes + 1;
    vfrb_status st = VFRB_OK;
    hipError_t e = hipGetDevice(&S->device);
    if (e == hipSuccess) st = vfrb_fb_plan_create(&S->fbp, d->twonchan, d->nave, d->nskip, d->nchans);
    if (e == hipSuccess && st == VFRB_OK) e = hipMalloc(&S->arena_base, (size_t)d->arena_bytes);
    S->arena.reset(d->arena_bytes);
    const int nlanes = d->batched ? 2 : 1;
    if (e == hipSuccess && st == VFRB_OK && nlanes > 1) e = hipStreamCreateWithFlags(&S->copy_s, hipStreamNonBlocking);
    S->lanes.resize(nlanes);
    for (Lane& l : S->lanes) {
        if (e != hipSuccess || st != VFRB_OK) break;
        e = hipStreamCreateWithFlags(&l.s, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&l.ev0);
        if (e == hipSuccess) e = hipEventCreate(&l.ev1);
        if (e == hipSuccess && nlanes > 1) e = hipEventCreateWithFlags(&l.ev_copy, hipEventDisableTiming);
        l.batches.assign(vfrb_search_tile_count(), nullptr);
        if (e == hipSuccess && d->batched) {
            // the batch of the configured (or default) tile up front; others on first use
            vfrb_batch* b = nullptr;
            st = vfrb_batch_create(&b, d->storage, d->tile_config, d->max_streams, 1 << 22);
            if (st == VFRB_OK) l.batches[vfrb_batch_get_tile(b)] = b;
        }
    }
    if (e != hipSuccess) st = vfrb::hip_fail(e, "vfrb_server_create", __FILE__, __LINE__);
    if (st != VFRB_OK) {
        server_free(S);
        return st;
    }
    S->sched = std::thread(scheduler_loop, S);
    S->compl_ = std::thread(completer_loop, S);
    *server = S;
    return VFRB_OK;
}

vfrb_status vfrb_server_destroy(vfrb_server* S) {
    if (!S) return VFRB_OK;
    vfrb_server_drain(S);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_sched.notify_all();
    S->cv_done.notify_all();
    S->sched.join();
    S->compl_.join();
    server_free(S);
    return VFRB_OK;
}

// FNV-1a over the values that give a stream's saved state its meaning: the
// filterbank shape and the delays (band, sample time, DM list).
static uint64_t stream_key(const vfrb_server_desc& d, const float* dms, int32_t ndm) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
    };
    const int32_t shape[4] = {d.twonchan, d.nave, d.nchans, d.nskip};
    const double band[3] = {d.fch1, d.chanbw, d.tsamp};
    mix(shape, sizeof shape);
    mix(band, sizeof band);
    mix(&ndm, sizeof ndm);
    mix(dms, sizeof(float) * (size_t)ndm);
    return h;
}

vfrb_status vfrb_server_open_stream(vfrb_server* S, const float* dms, int32_t ndm, int32_t* stream_id) {
    VFRB_REQUIRE(S && dms && ndm > 0 && stream_id, VFRB_ERR_INVALID, "bad arguments");
    {
        std::lock_guard<std::mutex> lk(S->mu);
        VFRB_REQUIRE((int)S->streams.size() < S->d.max_streams, VFRB_ERR_CAPACITY,
                     "server holds %d streams already", S->d.max_streams);
    }
    const vfrb_server_desc& d = S->d;
    std::vector<int32_t> off((size_t)ndm * d.nchans);
    int32_t max_delay = 0;
    vfrb_status st = vfrb_delay_table(dms, ndm, d.nchans, d.fch1, d.chanbw, d.tsamp, off.data(),
                                      &max_delay, nullptr);
    if (st != VFRB_OK) return st;
    auto ss = std::make_unique<SrvStream>();
    SrvStream* s = ss.get();
    s->ndm = ndm;
    s->max_delay = max_delay;
    s->key = stream_key(d, dms, ndm);
    s->ring_rows = max_delay + S->max_seg_rows + 1;
    s->ld_out = (S->max_seg_rows + 1 + 63) / 64 * 64;
    s->lay.set(ndm, d.max_cand);
    // admission of the stream itself: its resident state must fit in free HBM
    const size_t esz = d.storage == VFRB_STORE_F32 ? 4 : d.storage == VFRB_STORE_F16 ? 2 : 1;
    const size_t need = (size_t)s->ring_rows * d.nchans * 4 + (size_t)(s->ring_rows + 64) * d.nchans * esz +
                        (size_t)ndm * s->ld_out * 4 + (size_t)ndm * d.nchans * 4 + s->lay.bytes +
                        2 * (size_t)S->row_bytes;
    size_t free_b = 0, total_b = 0;
    VFRB_HIP(hipMemGetInfo(&free_b, &total_b));
    VFRB_REQUIRE(need + (size_t)(256u << 20) <= free_b, VFRB_ERR_NOMEM,
                 "stream state needs %zu bytes, %zu free", need, free_b);
    vfrb_search_desc sd{};
    sd.nchans = d.nchans;
    sd.tsamp = d.tsamp;
    sd.fch1 = d.fch1;
    sd.chanbw = d.chanbw;
    sd.ndm = ndm;
    sd.dms = dms;
    sd.max_rows = (int32_t)s->ring_rows;
    sd.storage = d.storage;
    sd.tile_config = d.tile_config;
    st = vfrb_search_plan_create(&s->plan, &sd);
    hipError_t e = hipSuccess;
    if (st == VFRB_OK) e = hipMalloc(&s->ring, (size_t)s->ring_rows * d.nchans * sizeof(float));
    if (st == VFRB_OK && e == hipSuccess) e = hipMalloc(&s->carry, 2 * (size_t)S->row_bytes);
    if (st == VFRB_OK && e == hipSuccess) e = hipMalloc(&s->out, (size_t)ndm * s->ld_out * sizeof(float));
    if (st == VFRB_OK && e == hipSuccess) e = hipMalloc(&s->d_res, s->lay.bytes);
    if (st == VFRB_OK && e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->h_res), s->lay.bytes);
    if (st == VFRB_OK && e != hipSuccess) st = vfrb::hip_fail(e, "vfrb_server_open_stream", __FILE__, __LINE__);
    if (st != VFRB_OK) {
        free_stream(s);
        return st;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    VFRB_REQUIRE((int)S->streams.size() < d.max_streams, VFRB_ERR_CAPACITY, "stream capacity reached");
    s->id = (int)S->streams.size();
    *stream_id = s->id;
    S->streams.push_back(std::move(ss));
    return VFRB_OK;
}

int32_t vfrb_server_max_delay(const vfrb_server* S, int32_t id) {
    if (!S) return -1;
    std::lock_guard<std::mutex> lk(const_cast<vfrb_server*>(S)->mu);
    if (id < 0 || id >= (int)S->streams.size()) return -1;
    return S->streams[id]->max_delay;
}

static vfrb_status submit(vfrb_server* S, int32_t stream_id, const uint8_t* pol0, const uint8_t* pol1,
                          int64_t nbytes, int64_t tag, const vfrb_server_output* out, bool device) {
    VFRB_REQUIRE(S && pol0 && pol1 && nbytes > 0, VFRB_ERR_INVALID, "bad segment");
    VFRB_REQUIRE(nbytes <= S->d.max_segment_bytes, VFRB_ERR_CAPACITY,
                 "segment of %lld bytes exceeds max_segment_bytes %lld", (long long)nbytes,
                 (long long)S->d.max_segment_bytes);
    VFRB_REQUIRE(2 * align_up(nbytes + S->row_bytes) <= S->d.arena_bytes, VFRB_ERR_CAPACITY,
                 "segment can never fit the input arena");
    Request r;
    r.tag = tag;
    r.stream = stream_id;
    r.pol[0] = pol0;
    r.pol[1] = pol1;
    r.nbytes = nbytes;
    r.device = device;
    if (out) r.out = *out;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        VFRB_REQUIRE(stream_id >= 0 && stream_id < (int)S->streams.size(), VFRB_ERR_INVALID,
                     "stream %d not open", stream_id);
        r.arrival = S->now();
        S->pending.push_back(r);
        S->submitted++;
    }
    S->cv_sched.notify_one();
    return VFRB_OK;
}

vfrb_status vfrb_server_submit(vfrb_server* S, int32_t stream_id, const uint8_t* pol0,
                               const uint8_t* pol1, int64_t nbytes, int64_t tag,
                               const vfrb_server_output* out) {
    return submit(S, stream_id, pol0, pol1, nbytes, tag, out, false);
}

vfrb_status vfrb_server_submit_device(vfrb_server* S, int32_t stream_id, const uint8_t* pol0,
                                      const uint8_t* pol1, int64_t nbytes, int64_t tag,
                                      const vfrb_server_output* out) {
    return submit(S, stream_id, pol0, pol1, nbytes, tag, out, true);
}

vfrb_status vfrb_server_poll(vfrb_server* S, vfrb_server_result* results, int32_t max_results,
                             int32_t wait_ms, int32_t* n) {
    VFRB_REQUIRE(S && n && (results || max_results == 0), VFRB_ERR_INVALID, "bad arguments");
    std::unique_lock<std::mutex> lk(S->mu);
    // something to return, or nothing outstanding that could produce it
    auto ready = [&] { return !S->completed.empty() || S->finished >= S->submitted; };
    if (S->completed.empty() && wait_ms != 0) {
        if (wait_ms < 0)
            S->cv_done.wait(lk, ready);
        else
            S->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
    }
    int k = 0;
    while (k < max_results && !S->completed.empty()) {
        results[k++] = S->completed.front();
        S->completed.pop_front();
    }
    *n = k;
    return VFRB_OK;
}

vfrb_status vfrb_server_drain(vfrb_server* S) {
    VFRB_REQUIRE(S, VFRB_ERR_INVALID, "NULL server");
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return S->finished >= S->submitted; });
    return VFRB_OK;
}

vfrb_status vfrb_server_get_stats(const vfrb_server* S, vfrb_server_stats* out) {
    VFRB_REQUIRE(S && out, VFRB_ERR_INVALID, "NULL argument");
    vfrb_server* m = const_cast<vfrb_server*>(S);
    std::lock_guard<std::mutex> lk(m->mu);
    out->submitted = S->submitted;
    out->completed = S->finished;
    out->iterations = S->iterations;
    out->batched_segments = S->batched_segments;
    out->arena_peak_bytes = m->arena.peak();
    out->device_busy_s = S->busy_s;
    return VFRB_OK;
}

double vfrb_server_now(const vfrb_server* S) { return S ? S->now() : 0.0; }

}  // extern "C"

// ---------------------------------------------------------------------------
// Stream state: the overlap rows (the raw filterbank rows the next window
// starts with, f32), the payload bytes short of a row, and the row count.
// ---------------------------------------------------------------------------
namespace {

struct StreamStateHeader {
    char magic[8];
    int32_t version, nchans, ndm, max_delay;
    uint64_t key;
    int64_t total_rows, keep_rows, carry_bytes, row_bytes;
};
constexpr char kStateMagic[8] = {'V', 'F', 'R', 'B', 'S', 'T', 'R', 'M'};
constexpr int32_t kStateVersion = 1;

int64_t keep_rows(const SrvStream* st) { return std::min<int64_t>(st->total_rows, st->max_delay); }

int64_t state_bytes(const vfrb_server* S, int64_t keep, int64_t carry) {
    return (int64_t)sizeof(StreamStateHeader) + (2 * carry + 3) / 4 * 4 + keep * S->d.nchans * 4;
}

// Waits (lock held) until stream id has no segment queued or in flight, then
// marks it busy so that the scheduler leaves it alone while its state moves.
vfrb_status claim_idle(vfrb_server* S, int32_t id, SrvStream** out) {
    std::unique_lock<std::mutex> lk(S->mu);
    VFRB_REQUIRE(id >= 0 && id < (int)S->streams.size(), VFRB_ERR_INVALID, "stream %d not open", id);
    SrvStream* st = S->streams[id].get();
    S->cv_done.wait(lk, [&] {
        if (st->busy) return false;
        for (const Request& r : S->pending)
            if (r.stream == id) return false;
        return true;
    });
    st->busy = true;
    *out = st;
    return VFRB_OK;
}

void release(vfrb_server* S, SrvStream* st) {
    {
        std::lock_guard<std::mutex> lk(S->mu);
        st->busy = false;
    }
    S->cv_sched.notify_one();
    S->cv_done.notify_all();
}

// Copies rows [r0, r0 + n) of a stream's ring (row r at r mod ring_rows)
// to or from host rows, in at most two contiguous pieces.
hipError_t ring_copy(SrvStream* st, int nch, int64_t r0, int64_t n, float* host, bool to_host) {
    int64_t done = 0;
    while (done < n) {
        const int64_t pos = (r0 + done) % st->ring_rows;
        const int64_t len = std::min(n - done, st->ring_rows - pos);
        float* dev = st->ring + pos * nch;
        float* h = host + done * nch;
        const size_t bytes = (size_t)len * nch * sizeof(float);
        const hipError_t e = to_host ? hipMemcpy(h, dev, bytes, hipMemcpyDeviceToHost)
                                     : hipMemcpy(dev, h, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) return e;
        done += len;
    }
    return hipSuccess;
}

}  // namespace

extern "C" {

vfrb_status vfrb_server_save_stream(vfrb_server* S, int32_t stream_id, void* buf, int64_t cap, int64_t* nbytes) {
    VFRB_REQUIRE(S && nbytes, VFRB_ERR_INVALID, "NULL argument");
    SrvStream* st = nullptr;
    vfrb_status rc = claim_idle(S, stream_id, &st);
    if (rc != VFRB_OK) return rc;
    const int64_t keep = keep_rows(st), carry = st->carry_bytes;
    const int64_t need = state_bytes(S, keep, carry);
    *nbytes = need;
    if (buf && cap < need) {
        release(S, st);
        VFRB_REQUIRE(false, VFRB_ERR_CAPACITY, "stream state needs %lld bytes, buffer holds %lld",
                     (long long)need, (long long)cap);
    }
    if (buf) {
        StreamStateHeader h{};
        std::memcpy(h.magic, kStateMagic, sizeof h.magic);
        h.version = kStateVersion;
        h.nchans = S->d.nchans;
        h.ndm = st->ndm;
        h.max_delay = st->max_delay;
        h.key = st->key;
        h.total_rows = st->total_rows;
        h.keep_rows = keep;
        h.carry_bytes = carry;
        h.row_bytes = S->row_bytes;
        unsigned char* b = static_cast<unsigned char*>(buf);
        std::memcpy(b, &h, sizeof h);
        unsigned char* cb = b + sizeof h;
        hipError_t e = hipSuccess;
        for (int p = 0; p < 2 && e == hipSuccess && carry > 0; p++)
            e = hipMemcpy(cb + p * carry, st->carry + p * S->row_bytes, (size_t)carry, hipMemcpyDeviceToHost);
        float* rows = reinterpret_cast<float*>(cb + (2 * carry + 3) / 4 * 4);
        if (e == hipSuccess) e = ring_copy(st, S->d.nchans, st->total_rows - keep, keep, rows, true);
        if (e != hipSuccess) rc = vfrb::hip_fail(e, "vfrb_server_save_stream", __FILE__, __LINE__);
    }
    release(S, st);
    return rc;
}

vfrb_status vfrb_server_restore_stream(vfrb_server* S, int32_t stream_id, const void* buf, int64_t nbytes) {
    VFRB_REQUIRE(S && buf, VFRB_ERR_INVALID, "NULL argument");
    VFRB_REQUIRE(nbytes >= (int64_t)sizeof(StreamStateHeader), VFRB_ERR_INVALID,
                 "%lld bytes hold no stream state", (long long)nbytes);
    StreamStateHeader h;
    std::memcpy(&h, buf, sizeof h);
    VFRB_REQUIRE(std::memcmp(h.magic, kStateMagic, sizeof h.magic) == 0 && h.version == kStateVersion,
                 VFRB_ERR_INVALID, "not a stream state of this version");
    VFRB_REQUIRE(h.total_rows >= 0 && h.keep_rows == std::min<int64_t>(h.total_rows, h.max_delay) &&
                     h.carry_bytes >= 0 && h.carry_bytes < h.row_bytes,
                     VFRB_ERR_INVALID, "inconsistent stream state");
    VFRB_REQUIRE(nbytes == state_bytes(S, h.keep_rows, h.carry_bytes) && h.nchans == S->d.nchans &&
                     h.row_bytes == S->row_bytes,
                 VFRB_ERR_INVALID, "stream state of another filterbank shape or truncated");
    SrvStream* st = nullptr;
    vfrb_status rc = claim_idle(S, stream_id, &st);
    if (rc != VFRB_OK) return rc;
    if (h.key != st->key || h.ndm != st->ndm || h.max_delay != st->max_delay) {
        release(S, st);
        VFRB_REQUIRE(false, VFRB_ERR_INVALID, "stream state of another band or DM list");
    }
    const unsigned char* cb = static_cast<const unsigned char*>(buf) + sizeof h;
    hipError_t e = hipSuccess;
    for (int p = 0; p < 2 && e == hipSuccess && h.carry_bytes > 0; p++)
        e = hipMemcpy(st->carry + p * S->row_bytes, cb + p * h.carry_bytes, (size_t)h.carry_bytes,
                      hipMemcpyHostToDevice);
    const float* rows = reinterpret_cast<const float*>(cb + (2 * h.carry_bytes + 3) / 4 * 4);
    if (e == hipSuccess)
        e = ring_copy(st, S->d.nchans, h.total_rows - h.keep_rows, h.keep_rows, const_cast<float*>(rows), false);
    if (e == hipSuccess) {
        st->total_rows = h.total_rows;
        st->carry_bytes = h.carry_bytes;
    } else {
        rc = vfrb::hip_fail(e, "vfrb_server_restore_stream", __FILE__, __LINE__);
    }
    release(S, st);
    return rc;
}

}  // extern "C"
