// Serving loop: render jobs of mixed frame shapes arrive at any time and run
// concurrently on one GPU.  Each running job owns a hider context (its own
// HIP stream), so jobs overlap on the device; contexts stay resident after a
// job, keyed by frame shape, so a later job of the same shape reuses the
// sample plan and device buffers.  Submission only copies the inputs: a pool
// of scheduler threads computes each job's split/dice schedule (which fixes
// its device footprint), and a job is admitted, in the order its schedule
// completed, when a slot is free and the device memory of all resident
// contexts plus its footprint fits the budget (idle contexts are evicted
// oldest first to make room).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/aqsis_hider.h"
#include "front.h"

uint64_t rh_internal_device_bytes(const rh_ctx* c);
void rh_internal_trim_pool(uint64_t keep);
bool rh_internal_schedule(const rh_options& opt, const rh_prim* prims, int32_t n, int threads, rh::Schedule& S,
                          rh::ForestPtr& forest, uint64_t& bytes, double& ms, std::string& err);
rh_status rh_internal_post_scheduled(rh_ctx* c, const rh_prim* prims, int32_t n, rh::Schedule&& S,
                                     rh::ForestPtr&& forest, double ms);

namespace {

using Clock = std::chrono::steady_clock;

thread_local std::string s_err;

rh_status sfail(rh_status code, const std::string& msg) {
    s_err = msg;
    return code;
}

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Frame shape: the options that fix a context's sample plan and buffers.
std::string shape_key(const rh_options& o) { return std::string(reinterpret_cast<const char*>(&o), sizeof(o)); }

struct Resident {
    rh_ctx* ctx = nullptr;
    std::string key;
    uint64_t bytes = 0;                   // device memory it holds
    bool busy = false;
    Clock::time_point lastUse;
};

struct Job {
    uint64_t id = 0;
    rh_options opt;
    std::vector<rh_prim> prims;
    uint64_t estimate = 0;                // device bytes of the scheduled frame
    rh::Schedule sched;                   // computed at submission, used once
    rh::ForestPtr forest;                 // its split trees (for the culling schedule)
    double msSchedule = 0.0;
    int32_t status = RH_OK;
    int state = 0;                        // 0 queued, 1 running, 2 done
    bool collected = false;
    std::string error;
    int32_t width = 0, height = 0;
    std::vector<float> channels;
    std::vector<uint8_t> rgba8;
    rh_stats stats;
    double msSetup = 0.0, msRender = 0.0, msReadback = 0.0;
    Clock::time_point tSubmit, tReady, tStart, tEnd;
};

}  // namespace

// Backfill limits of the admission queue (see worker()).
constexpr double kBackfillMs = 1000.0;
constexpr int kBackfillDepth = 16;

struct rh_server {
    rh_server_options opt;
    std::mutex m;
    std::condition_variable cvWork, cvDone, cvSched;
    std::deque<Job*> schedQueue;          // submitted, schedule not yet computed
    std::deque<Job*> queue;               // scheduled, waiting for admission
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    std::list<Resident> resident;
    uint64_t nextId = 1;
    uint64_t maxBytes = 0;                // high-water mark of resident device memory
    bool stop = false;
    int threadsPerSchedule = 1;
    std::vector<std::thread> workers, schedulers;

    // Under the lock: a job ends without running.
    void finish_unrun(Job* j, int32_t status, const std::string& msg) {
        j->status = status;
        j->error = msg;
        j->state = 2;
        j->tStart = j->tEnd = Clock::now();
        j->prims.clear();
        j->prims.shrink_to_fit();
    }

    void scheduler() {
        for (;;) {
            Job* j = nullptr;
            {
                std::unique_lock<std::mutex> lk(m);
                cvSched.wait(lk, [&] { return stop || !schedQueue.empty(); });
                if (stop) {
                    while (!schedQueue.empty()) {
                        finish_unrun(schedQueue.front(), RH_ERR_STATE, "server stopped");
                        schedQueue.pop_front();
                    }
                    cvDone.notify_all();
                    return;
                }
                j = schedQueue.front();
                schedQueue.pop_front();
            }
            std::string err;
            const bool ok = rh_internal_schedule(j->opt, j->prims.data(), (int32_t)j->prims.size(),
                                                 threadsPerSchedule, j->sched, j->forest, j->estimate, j->msSchedule,
                                                 err);
            {
                std::lock_guard<std::mutex> lk(m);
                if (!ok) {
                    finish_unrun(j, RH_ERR_INVALID, err);
                } else if (j->estimate > opt.memory_budget) {
                    finish_unrun(j, RH_ERR_CAPACITY, "job needs " + std::to_string(j->estimate) +
                                                         " device bytes, more than the server's budget");
                } else {
                    j->tReady = Clock::now();
                    queue.push_back(j);
                }
            }
            cvDone.notify_all();
            cvWork.notify_all();
        }
    }

    // Evicted contexts wait here to be destroyed outside the lock (freeing
    // device memory can take a while); their bytes count until then.
    std::vector<std::pair<rh_ctx*, uint64_t>> graveyard;
    uint64_t pendingBytes = 0;

    uint64_t liveBytes() const {
        uint64_t t = 0;
        for (const auto& r : resident) t += r.bytes;
        return t;
    }
    uint64_t residentBytes() const { return liveBytes() + pendingBytes; }

    // Not under the lock: destroy what was evicted.
    void bury() {
        std::vector<std::pair<rh_ctx*, uint64_t>> dead;
        {
            std::lock_guard<std::mutex> lk(m);
            dead.swap(graveyard);
        }
        if (dead.empty()) return;
        for (auto& d : dead) rh_destroy(d.first);
        // the pool holds what the dead contexts freed: keep only the budget
        rh_internal_trim_pool(opt.memory_budget);
        {
            std::lock_guard<std::mutex> lk(m);
            for (auto& d : dead) pendingBytes -= d.second;
        }
        cvWork.notify_all();
    }

    // Under the lock: pick a context for the head job if it can be admitted.
    bool admit(Job* j, std::list<Resident>::iterator& slot) {
        const std::string key = shape_key(j->opt);
        // an idle resident context of the same shape
        for (auto it = resident.begin(); it != resident.end(); ++it)
            if (!it->busy && it->key == key) {
                const uint64_t extra = j->estimate > it->bytes ? j->estimate - it->bytes : 0;
                if (residentBytes() + extra > opt.memory_budget && !evict(extra, &*it)) return false;
                it->busy = true;
                it->bytes = std::max(it->bytes, j->estimate);
                maxBytes = std::max(maxBytes, residentBytes());
                slot = it;
                return true;
            }
        if (residentBytes() + j->estimate > opt.memory_budget && !evict(j->estimate, nullptr)) return false;
        maxBytes = std::max(maxBytes, residentBytes() + j->estimate);
        Resident r;
        r.key = key;
        r.bytes = j->estimate;
        r.busy = true;
        resident.push_back(r);
        slot = std::prev(resident.end());
        return true;
    }

    // Evict idle contexts, oldest first, until `need` more bytes fit once
    // the evicted ones are destroyed; true if they fit now.
    bool evict(uint64_t need, const Resident* keep) {
        while (liveBytes() + need > opt.memory_budget) {
            auto victim = resident.end();
            for (auto it = resident.begin(); it != resident.end(); ++it)
                if (!it->busy && &*it != keep && (victim == resident.end() || it->lastUse < victim->lastUse))
                    victim = it;
            if (victim == resident.end()) return false;
            if (victim->ctx) {
                graveyard.emplace_back(victim->ctx, victim->bytes);
                pendingBytes += victim->bytes;
            }
            resident.erase(victim);
        }
        return residentBytes() + need <= opt.memory_budget;
    }

    void run(Job* j, Resident& r) {
        rh_status st = RH_OK;
        const Clock::time_point t0 = Clock::now();
        if (!r.ctx) st = rh_create(&j->opt, &r.ctx);
        const Clock::time_point t1 = Clock::now();
        if (st == RH_OK)
            st = rh_internal_post_scheduled(r.ctx, j->prims.data(), (int32_t)j->prims.size(), std::move(j->sched),
                                            std::move(j->forest), j->msSchedule);
        if (st == RH_OK) st = rh_render(r.ctx);
        const Clock::time_point t2 = Clock::now();
        if (st == RH_OK) st = rh_frame_size(r.ctx, &j->width, &j->height, nullptr, nullptr, nullptr);
        if (st == RH_OK) {
            j->channels.resize((size_t)j->width * j->height * 9);
            j->rgba8.resize((size_t)j->width * j->height * 4);
            st = rh_read_channels(r.ctx, j->channels.data());
        }
        if (st == RH_OK) st = rh_read_rgba8(r.ctx, j->rgba8.data());
        j->msSetup = ms_between(t0, t1);
        j->msRender = ms_between(t1, t2);
        j->msReadback = ms_between(t2, Clock::now());
        if (st == RH_OK) st = rh_get_stats(r.ctx, &j->stats);
        if (st != RH_OK) j->error = rh_last_error();
        j->status = st;
        j->prims.clear();
        j->prims.shrink_to_fit();
    }

    void worker() {
        for (;;) {
            Job* j = nullptr;
            std::list<Resident>::iterator slot;
            {
                std::unique_lock<std::mutex> lk(m);
                for (;;) {
                    if (stop) break;
                    if (!queue.empty()) {
                        // The head job first.  While it has waited less than
                        // kBackfillMs for room, a later job that fits may go
                        // ahead of it (a large frame waiting for memory does
                        // not hold back small ones); after that, strict order,
                        // so the head is never starved.
                        auto pick = queue.end();
                        if (admit(queue.front(), slot)) {
                            pick = queue.begin();
                        } else if (ms_between(queue.front()->tReady, Clock::now()) < kBackfillMs) {
                            int tried = 0;
                            for (auto it = std::next(queue.begin()); it != queue.end() && tried < kBackfillDepth;
                                 ++it, ++tried)
                                if (admit(*it, slot)) {
                                    pick = it;
                                    break;
                                }
                        }
                        if (pick != queue.end()) {
                            j = *pick;
                            queue.erase(pick);
                            break;
                        }
                    }
                    if (!graveyard.empty()) {
                        // room is being made: destroy the evicted contexts, retry
                        lk.unlock();
                        bury();
                        lk.lock();
                        continue;
                    }
                    cvWork.wait_for(lk, std::chrono::milliseconds(50));
                }
                if (!j) {
                    // stopped: fail what was never admitted
                    while (!queue.empty()) {
                        finish_unrun(queue.front(), RH_ERR_STATE, "server stopped");
                        queue.pop_front();
                    }
                    cvDone.notify_all();
                    return;
                }
                j->state = 1;
                j->tStart = Clock::now();
            }
            cvWork.notify_all();          // the next job may fit as well
            bury();
            run(j, *slot);
            {
                std::lock_guard<std::mutex> lk(m);
                if (j->status != RH_OK || !slot->ctx) {
                    if (slot->ctx) {
                        graveyard.emplace_back(slot->ctx, slot->bytes);
                        pendingBytes += slot->bytes;
                    }
                    resident.erase(slot);
                } else {
                    slot->bytes = rh_internal_device_bytes(slot->ctx);
                    slot->busy = false;
                    slot->lastUse = Clock::now();
                }
                maxBytes = std::max(maxBytes, residentBytes());
                // a frame may have grown past its estimate (a transparent-hit
                // arena redone at its exact size): back within the budget,
                // dropping idle contexts oldest first
                evict(0, nullptr);
                j->tEnd = Clock::now();
                j->state = 2;
            }
            // blocks the job's buffers outgrew stay in the pool: keep the
            // pool within the budget
            rh_internal_trim_pool(opt.memory_budget);
            cvDone.notify_all();
            cvWork.notify_all();
            bury();
        }
    }

    void fill(const Job& j, rh_job_info* out) const {
        std::memset(out, 0, sizeof(*out));
        out->id = j.id;
        out->status = j.status;
        out->state = j.state;
        out->width = j.width;
        out->height = j.height;
        out->device_bytes = j.estimate;
        if (j.state >= 1) out->ms_queue = ms_between(j.tSubmit, j.tStart);
        if (j.state == 2) {
            out->ms_run = ms_between(j.tStart, j.tEnd);
            out->ms_total = ms_between(j.tSubmit, j.tEnd);
            out->ms_device = j.stats.ms_device;
            out->spl_count = j.stats.spl_count;
            out->ms_setup = j.msSetup;
            out->ms_upload = j.stats.ms_upload;
            out->ms_render = j.msRender;
            out->ms_readback = j.msReadback;
        }
    }
};

extern "C" {

rh_status rh_server_create(const rh_server_options* opt, rh_server** out) {
    if (!opt || !out) return sfail(RH_ERR_INVALID, "null argument");
    *out = nullptr;
    if (opt->slots < 1 || opt->slots > 64) return sfail(RH_ERR_INVALID, "slots must be in 1..64");
    if (opt->memory_budget == 0) return sfail(RH_ERR_INVALID, "memory budget must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return sfail(RH_ERR_DEVICE, "no HIP device available");
    rh_server* s = new (std::nothrow) rh_server();
    if (!s) return sfail(RH_ERR_NOMEM, "out of host memory");
    s->opt = *opt;
    if (s->opt.schedule_threads <= 0) s->opt.schedule_threads = 16;
    // schedulers of about four threads each
    const int nsched = std::max(1, s->opt.schedule_threads / 4);
    s->threadsPerSchedule = std::max(1, s->opt.schedule_threads / nsched);
    for (int i = 0; i < nsched; ++i) s->schedulers.emplace_back([s] { s->scheduler(); });
    for (int i = 0; i < s->opt.slots; ++i) s->workers.emplace_back([s] { s->worker(); });
    *out = s;
    return RH_OK;
}

rh_status rh_server_submit(rh_server* s, const rh_options* opt, const rh_prim* prims, int32_t n, uint64_t* id) {
    if (!s || !opt || (n > 0 && !prims) || n < 0 || !id) return sfail(RH_ERR_INVALID, "bad arguments");
    std::unique_ptr<Job> j(new Job());
    j->opt = *opt;
    j->prims.assign(prims, prims + n);
    j->tSubmit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (s->stop) return sfail(RH_ERR_STATE, "server stopped");
        j->id = s->nextId++;
        *id = j->id;
        s->schedQueue.push_back(j.get());
        s->jobs[j->id] = std::move(j);
    }
    s->cvSched.notify_one();
    return RH_OK;
}

rh_status rh_server_wait(rh_server* s, uint64_t id, rh_job_info* info) {
    if (!s) return sfail(RH_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(RH_ERR_INVALID, "unknown job");
    Job* j = it->second.get();
    s->cvDone.wait(lk, [&] { return j->state == 2; });
    j->collected = true;
    if (info) s->fill(*j, info);
    if (j->status != RH_OK) return sfail(j->status, j->error);
    return RH_OK;
}

rh_status rh_server_poll(rh_server* s, rh_job_info* infos, int32_t max, int32_t* count) {
    if (!s || !count || max < 0 || (max > 0 && !infos)) return sfail(RH_ERR_INVALID, "bad arguments");
    std::lock_guard<std::mutex> lk(s->m);
    int32_t k = 0;
    for (auto& kv : s->jobs) {
        if (k >= max) break;
        Job* j = kv.second.get();
        if (j->state == 2 && !j->collected) {
            j->collected = true;
            s->fill(*j, &infos[k++]);
        }
    }
    *count = k;
    return RH_OK;
}

rh_status rh_server_read(rh_server* s, uint64_t id, float* channels, uint8_t* rgba8) {
    if (!s) return sfail(RH_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(RH_ERR_INVALID, "unknown job");
    Job* j = it->second.get();
    if (j->state != 2) return sfail(RH_ERR_STATE, "job not finished");
    if (j->status != RH_OK) return sfail(j->status, j->error);
    if (channels) std::memcpy(channels, j->channels.data(), j->channels.size() * sizeof(float));
    if (rgba8) std::memcpy(rgba8, j->rgba8.data(), j->rgba8.size());
    return RH_OK;
}

rh_status rh_server_release(rh_server* s, uint64_t id) {
    if (!s) return sfail(RH_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(RH_ERR_INVALID, "unknown job");
    if (it->second->state != 2) return sfail(RH_ERR_STATE, "job not finished");
    s->jobs.erase(it);
    return RH_OK;
}

rh_status rh_server_memory(rh_server* s, uint64_t* resident, uint64_t* high_water) {
    if (!s) return sfail(RH_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->m);
    if (resident) *resident = s->residentBytes();
    if (high_water) *high_water = s->maxBytes;
    return RH_OK;
}

void rh_server_destroy(rh_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cvWork.notify_all();
    s->cvSched.notify_all();
    for (auto& t : s->schedulers) t.join();
    for (auto& t : s->workers) t.join();
    s->bury();
    for (auto& r : s->resident)
        if (r.ctx) rh_destroy(r.ctx);
    delete s;
}

const char* rh_server_last_error(void) { return s_err.c_str(); }

}  // extern "C"
