// Device-resident commit server: continuous admission of LDE + commit jobs.
//
// The reference serves proofs by forking one prover process per request and
// keeps nothing on the device between jobs (src/prover_server/prover.cpp:
// 459-505; rust/prover_server/src/server.rs:22, 70-95 admit one job per GPU).
// Here one process owns the GPU: jobs queue FIFO; a scheduler thread admits
// the head job as soon as a stream slot is free and its buffers fit in the
// HBM arena (first-fit sub-allocation of one allocation made at start-up, so
// admission never calls hipMalloc/hipFree and never synchronises the device);
// admitted jobs run on their own streams and overlap on the GPU.  A finished
// job keeps its codewords and tree resident for opening queries until it is
// released.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "arena.hpp"
#include "tvh_internal.hpp"

using namespace tvh;

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

constexpr size_t kAlign = 256;
inline size_t align_up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

// Byte layout of one job inside the arena: trace | rand | codewords | tree.
struct JobLayout {
    size_t trace = 0, rand = 0, cw = 0, tree = 0, total = 0;
    JobLayout() = default;
    explicit JobLayout(const tvh_job_desc& d) {
        size_t a = align_up(8 * d.trace_len * d.num_cols);
        size_t b = align_up(8 * d.rand_len * d.num_cols);
        size_t c = align_up(8 * d.target_len * d.num_cols);
        size_t t = align_up(8 * 5 * (2 * d.target_len - 1));
        rand = a;
        cw = a + b;
        tree = a + b + c;
        total = a + b + c + t;
    }
};

// out[q * w + c] = cw[c * m + rows[q]]  (codeword rows of a column-major table)
__global__ void k_gather_rows(const u64* __restrict__ cw, u64 m, u64 w,
                              const u64* __restrict__ rows, u64 count, u64* __restrict__ out) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count * w) return;
    u64 q = i / w, c = i - q * w;
    out[i] = cw[c * m + rows[q]];
}

}  // namespace

struct tvh_server {
    struct Job {
        tvh_job_desc d{};
        JobLayout lay{};
        int state = TVH_JOB_QUEUED;
        int status = TVH_OK;
        std::string err;
        size_t off = 0;
        hipEvent_t ev = nullptr;
        tvh_job_times t{};
        uint64_t root[5] = {0, 0, 0, 0, 0};
    };

    std::mutex mu;
    std::condition_variable cv;       // wakes the scheduler
    std::condition_variable done_cv;  // job completions
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Job> jobs;  // element references survive inserts
    std::vector<hipStream_t> slots;
    std::vector<int64_t> slot_job;
    u64* root_stage = nullptr;  // pinned host memory: 5 u64 per slot, written by the slot stream
    size_t waiters = 0;         // threads inside tvh_server_wait (destroy waits for them)
    std::condition_variable waiters_cv;
    tvh::FirstFitArena arena;
    uint8_t* base = nullptr;
    bool stop = false;
    uint64_t next_id = 1;
    size_t max_queue = 0;
    int device = 0;
    std::thread th;

    // opening queries: one query stream and a small buffer grown on demand
    std::mutex qmu;
    hipStream_t qstream = nullptr;
    u64* qbuf = nullptr;
    size_t qcap = 0;  // u64 elements

    u64* at(const Job& j, size_t part) { return (u64*)(base + j.off + part); }

    void finish(Job& j, int status, const char* msg) {
        j.state = status ? TVH_JOB_FAILED : TVH_JOB_DONE;
        j.status = status;
        if (status) j.err = msg ? msg : "job failed";
        j.t.done_s = now_s();
        done_cv.notify_all();
    }

    // Enqueue an admitted job on its slot stream: inputs, commit, then the
    // root into the slot's pinned staging words, then the completion event.
    int launch(Job& j, int slot) {
        hipStream_t st = slots[slot];
        const tvh_job_desc& d = j.d;
        u64* tr = at(j, j.lay.trace);
        u64* rd = at(j, j.lay.rand);
        size_t tn = d.trace_len * d.num_cols, rn = d.rand_len * d.num_cols;
        int rc = TVH_OK;
        if (d.host_trace) {
            rc = staged_h2d(tr, d.host_trace, 8 * tn, st);
        } else {
            rc = launch_fill_uniform(tr, tn, d.seed, st);
        }
        if (!rc && rn) {
            if (d.host_rand) {
                rc = staged_h2d(rd, d.host_rand, 8 * rn, st);
            } else {
                rc = launch_fill_uniform(rd, rn, d.seed ^ 0x5bd1e995ull, st);
            }
        }
        if (!rc)
            rc = tvh_commit_dev(tr, d.num_cols, d.trace_len, rn ? rd : nullptr, d.rand_len,
                                d.trace_offset, d.target_offset, d.target_len,
                                at(j, j.lay.cw), at(j, j.lay.tree), TVH_TREE_LEVEL_ORDER, st);
        if (!rc && hipMemcpyAsync(root_stage + 5 * slot, at(j, j.lay.tree) + 5 * (2 * d.target_len - 2),
                                  40, hipMemcpyDeviceToHost, st) != hipSuccess)
            rc = set_error(TVH_ERR_DEVICE, "root download");
        if (!rc && hipEventRecord(j.ev, st) != hipSuccess) rc = set_error(TVH_ERR_DEVICE, "event record");
        return rc;
    }

    void loop() {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        while (!stop) {
            bool progress = false;
            for (size_t s = 0; s < slots.size(); s++) {  // completions
                if (slot_job[s] < 0) continue;
                Job& j = jobs.at((uint64_t)slot_job[s]);
                hipError_t q = hipEventQuery(j.ev);
                if (q == hipErrorNotReady) continue;
                int rc = TVH_OK;
                if (q != hipSuccess)
                    rc = set_error(TVH_ERR_DEVICE, "job %lld: %s", (long long)slot_job[s],
                                   hipGetErrorString(q));
                else  // the event follows the root's copy into this slot's staging words
                    std::memcpy(j.root, root_stage + 5 * s, 40);
                if (rc) arena.release(j.off, j.lay.total);
                finish(j, rc, rc ? tvh_last_error() : nullptr);
                slot_job[s] = -1;
                progress = true;
            }
            while (!pending.empty()) {  // admissions, FIFO: the head waits for memory
                int slot = -1;
                for (size_t s = 0; s < slots.size() && slot < 0; s++)
                    if (slot_job[s] < 0) slot = (int)s;
                if (slot < 0) break;
                uint64_t id = pending.front();
                Job& j = jobs.at(id);
                if (j.lay.total > arena.size()) {
                    pending.pop_front();
                    int rc = set_error(TVH_ERR_OOM, "job %llu needs %zu bytes; the budget is %zu",
                                       (unsigned long long)id, j.lay.total, arena.size());
                    finish(j, rc, tvh_last_error());
                    progress = true;
                    continue;
                }
                size_t off;
                if (!arena.alloc(j.lay.total, &off)) break;
                pending.pop_front();
                j.off = off;
                j.t.admit_s = now_s();
                j.state = TVH_JOB_RUNNING;
                slot_job[slot] = (int64_t)id;  // the slot is taken while the lock is released
                // enqueue without the lock: a host-memory upload may take a while,
                // and waits, polls and opening queries of other jobs need the lock
                // (the job's map entry stays put: only finished jobs are erased)
                lk.unlock();
                int rc = launch(j, slot);
                lk.lock();
                if (rc) {
                    std::string msg = tvh_last_error();
                    // work already enqueued on the slot may still use the range
                    (void)hipStreamSynchronize(slots[slot]);
                    arena.release(j.off, j.lay.total);
                    finish(j, rc, msg.c_str());
                    slot_job[slot] = -1;
                }
                progress = true;
            }
            if (!progress) cv.wait_for(lk, std::chrono::microseconds(100));
        }
    }

    // Finished job by id (caller holds mu).
    Job* done_job(uint64_t id) {
        auto it = jobs.find(id);
        if (it == jobs.end() || it->second.state != TVH_JOB_DONE) return nullptr;
        return &it->second;
    }

    u64* query_buffer(size_t elems) {
        if (elems > qcap) {
            if (qbuf) (void)hipFree(qbuf);
            qbuf = nullptr;
            qcap = 0;
            size_t cap = std::max<size_t>(elems, 1 << 16);
            if (hipMalloc((void**)&qbuf, 8 * cap) != hipSuccess) return nullptr;
            qcap = cap;
        }
        return qbuf;
    }

    // Gathers tree nodes (heap indices) and, when rows_out is set, codeword rows.
    // The caller holds qmu (taken before the job lookup, so a concurrent release
    // of the job waits until the gather has finished).
    int gather(const u64* tree, const u64* cw, size_t m, size_t w, const std::vector<u64>& heap,
               const size_t* rows, size_t nrows, u64* nodes_out, u64* rows_out) {
        size_t nh = nodes_out ? heap.size() : 0, nr = rows_out ? nrows : 0;
        if (!nh && !(nr && w)) return TVH_OK;
        // buffer: heap indices | row indices | nodes (5 each) | rows (w each)
        size_t need = nh + nr + 5 * nh + nr * w;
        u64* b = query_buffer(need);
        if (!b) return set_error(TVH_ERR_OOM, "query buffer of %zu u64", need);
        u64 *d_idx = b, *d_rows = b + nh, *d_nodes = d_rows + nr, *d_rout = d_nodes + 5 * nh;
        std::vector<u64> host(nh + nr);
        std::copy(heap.begin(), heap.begin() + nh, host.begin());
        for (size_t q = 0; q < nr; q++) host[nh + q] = rows[q];
        int rc = TVH_OK;
        if (hipMemcpyAsync(b, host.data(), 8 * host.size(), hipMemcpyHostToDevice, qstream) != hipSuccess)
            rc = set_error(TVH_ERR_DEVICE, "query upload");
        if (!rc && nh) rc = launch_gather_heap(tree, m, TVH_TREE_LEVEL_ORDER, d_idx, nh, d_nodes, qstream);
        if (!rc && nr && w) {
            u64 total = nr * w;
            hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                               qstream, cw, (u64)m, (u64)w, (const u64*)d_rows, (u64)nr, d_rout);
            rc = check_launch();
        }
        if (!rc && nh &&
            hipMemcpyAsync(nodes_out, d_nodes, 40 * nh, hipMemcpyDeviceToHost, qstream) != hipSuccess)
            rc = set_error(TVH_ERR_DEVICE, "query download");
        if (!rc && nr && w &&
            hipMemcpyAsync(rows_out, d_rout, 8 * nr * w, hipMemcpyDeviceToHost, qstream) != hipSuccess)
            rc = set_error(TVH_ERR_DEVICE, "row download");
        if (hipStreamSynchronize(qstream) != hipSuccess && !rc) rc = set_error(TVH_ERR_DEVICE, "query sync");
        return rc;
    }
};

using Job = tvh_server::Job;

extern "C" {

int tvh_server_create(const tvh_server_config* cfg, tvh_server** out) {
    if (!out) return set_error(TVH_ERR_INVALID_ARG, "null server pointer");
    *out = nullptr;
    if (cfg && cfg->max_inflight < 0) return set_error(TVH_ERR_INVALID_ARG, "max_inflight < 0");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    size_t budget = cfg ? cfg->hbm_budget_bytes : 0;
    int nslots = cfg && cfg->max_inflight > 0 ? cfg->max_inflight : 4;
    if (budget == 0) {
        // memory freed just before may be released lazily: wait for the device
        // so that the free figure includes it
        if (hipDeviceSynchronize() != hipSuccess) return set_error(TVH_ERR_DEVICE, "device sync");
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return set_error(TVH_ERR_DEVICE, "hipMemGetInfo");
        budget = fr > (32ull << 30) ? fr - (16ull << 30) : fr / 2;
    }
    budget &= ~(kAlign - 1);
    if (!budget) return set_error(TVH_ERR_INVALID_ARG, "HBM budget below %zu bytes", kAlign);
    tvh_server* s = new tvh_server();
    s->device = ds->device;
    s->max_queue = cfg && cfg->max_queue ? cfg->max_queue : SIZE_MAX;
    if (hipMalloc((void**)&s->base, budget) != hipSuccess) {
        delete s;
        return set_error(TVH_ERR_OOM, "server arena of %zu bytes", budget);
    }
    s->arena.reset(budget);
    s->slots.assign(nslots, nullptr);
    s->slot_job.assign(nslots, -1);
    bool ok = hipHostMalloc((void**)&s->root_stage, 40 * (size_t)nslots, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&s->qstream, hipStreamNonBlocking) == hipSuccess;
    for (auto& st : s->slots) ok = ok && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
        for (auto st : s->slots)
            if (st) (void)hipStreamDestroy(st);
        if (s->qstream) (void)hipStreamDestroy(s->qstream);
        if (s->root_stage) (void)hipHostFree(s->root_stage);
        (void)hipFree(s->base);
        delete s;
        return set_error(TVH_ERR_DEVICE, "stream creation");
    }
    s->th = std::thread([s] { s->loop(); });
    *out = s;
    return TVH_OK;
}

int tvh_server_submit(tvh_server* s, const tvh_job_desc* d, uint64_t* job_id) {
    if (!s || !d || !job_id) return set_error(TVH_ERR_INVALID_ARG, "null argument");
    if (!is_pow2(d->trace_len) || !is_pow2(d->target_len))
        return set_error(TVH_ERR_NOT_POW2, "trace_len and target_len must be powers of 2");
    if (d->target_len < d->trace_len || d->rand_len > d->trace_len)
        return set_error(TVH_ERR_INVALID_ARG, "need rand_len <= trace_len <= target_len");
    // the job's coset tables are built here, outside the scheduler's lock, so
    // that admission finds them cached (no allocation, no wait)
    if (d->num_cols) {
        int rc = tvh_lde_prepare(d->trace_len, d->target_len, d->trace_offset, d->target_offset);
        if (rc) return rc;
    }
    hipEvent_t ev;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
        return set_error(TVH_ERR_DEVICE, "event");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop || s->pending.size() >= s->max_queue) {
        (void)hipEventDestroy(ev);
        return set_error(TVH_ERR_RANGE, s->stop ? "server is shutting down" : "queue full");
    }
    uint64_t id = s->next_id++;
    Job& j = s->jobs[id];
    j.d = *d;
    j.lay = JobLayout(*d);
    j.ev = ev;
    j.t.submit_s = now_s();
    s->pending.push_back(id);
    *job_id = id;
    s->cv.notify_all();
    return TVH_OK;
}

int tvh_server_poll(tvh_server* s, uint64_t id, int* state) {
    if (!s || !state) return set_error(TVH_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_error(TVH_ERR_RANGE, "unknown job %llu", (unsigned long long)id);
    *state = it->second.state;
    return TVH_OK;
}

int tvh_server_wait(tvh_server* s, uint64_t id, uint64_t* root, tvh_job_times* times) {
    if (!s) return set_error(TVH_ERR_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->jobs.find(id) == s->jobs.end())
        return set_error(TVH_ERR_RANGE, "unknown job %llu", (unsigned long long)id);
    // the job is looked up again after every wake-up: a concurrent release may
    // have erased it; destroy waits until no thread is left in here
    s->waiters++;
    Job* j = nullptr;
    s->done_cv.wait(lk, [&] {
        auto it = s->jobs.find(id);
        j = it == s->jobs.end() ? nullptr : &it->second;
        return !j || j->state >= TVH_JOB_DONE;
    });
    int rc = TVH_OK;
    if (!j) {
        rc = set_error(TVH_ERR_RANGE, "job %llu was released while waiting", (unsigned long long)id);
    } else {
        if (times) *times = j->t;
        if (j->state == TVH_JOB_FAILED) rc = set_error(j->status, "%s", j->err.c_str());
        else if (root) std::memcpy(root, j->root, 40);
    }
    if (--s->waiters == 0) s->waiters_cv.notify_all();
    return rc;
}

int tvh_server_open(tvh_server* s, uint64_t id, const size_t* leaves, size_t count,
                    uint64_t* rows_out, uint64_t* auth_out, size_t auth_cap, size_t* auth_count) {
    if (!s || (count && !leaves)) return set_error(TVH_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> q(s->qmu);  // lock order: qmu, then mu
    const u64 *tree, *cw;
    size_t m, w;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        Job* j = s->done_job(id);
        if (!j) return set_error(TVH_ERR_RANGE, "job %llu is not a finished job", (unsigned long long)id);
        m = j->d.target_len;
        w = j->d.num_cols;
        tree = s->at(*j, j->lay.tree);
        cw = s->at(*j, j->lay.cw);
    }
    size_t n = 0;
    int rc = tvh_auth_structure_indices(m, leaves, count, nullptr, 0, &n);
    if (rc) return rc;
    if (auth_count) *auth_count = n;
    if (auth_out && n > auth_cap)
        return set_error(TVH_ERR_INVALID_ARG, "auth capacity %zu < %zu", auth_cap, n);
    std::vector<u64> heap;
    if (auth_out) {
        heap.resize(n);
        rc = tvh_auth_structure_indices(m, leaves, count, heap.data(), n, &n);
        if (rc) return rc;
    }
    return s->gather(tree, cw, m, w, heap, leaves, count, auth_out, rows_out);
}

int tvh_server_paths(tvh_server* s, uint64_t id, const size_t* leaves, size_t count,
                     uint64_t* paths) {
    if (!s || (count && (!leaves || !paths))) return set_error(TVH_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> q(s->qmu);  // lock order: qmu, then mu
    const u64 *tree, *cw;
    size_t m, w;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        Job* j = s->done_job(id);
        if (!j) return set_error(TVH_ERR_RANGE, "job %llu is not a finished job", (unsigned long long)id);
        m = j->d.target_len;
        w = j->d.num_cols;
        tree = s->at(*j, j->lay.tree);
        cw = s->at(*j, j->lay.cw);
    }
    size_t h = log2u(m);
    std::vector<u64> heap(count * h);
    for (size_t q = 0; q < count; q++) {
        size_t c = 0;
        int rc = tvh_auth_path_indices(m, leaves[q], heap.data() + q * h, h, &c);
        if (rc) return rc;
    }
    return s->gather(tree, cw, m, w, heap, nullptr, 0, paths, nullptr);
}

int tvh_server_buffers(tvh_server* s, uint64_t id, const uint64_t** d_codewords,
                       const uint64_t** d_tree) {
    if (!s) return set_error(TVH_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    Job* j = s->done_job(id);
    if (!j) return set_error(TVH_ERR_RANGE, "job %llu is not a finished job", (unsigned long long)id);
    if (d_codewords) *d_codewords = s->at(*j, j->lay.cw);
    if (d_tree) *d_tree = s->at(*j, j->lay.tree);
    return TVH_OK;
}

int tvh_server_release(tvh_server* s, uint64_t id) {
    if (!s) return set_error(TVH_ERR_INVALID_ARG, "null server");
    // an opening query of this job may still be reading its range: take qmu first
    std::lock_guard<std::mutex> g(s->qmu);
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_error(TVH_ERR_RANGE, "unknown job %llu", (unsigned long long)id);
    Job& j = it->second;
    if (j.state < TVH_JOB_DONE)
        return set_error(TVH_ERR_INVALID_ARG, "job %llu is still active", (unsigned long long)id);
    if (j.state == TVH_JOB_DONE) s->arena.release(j.off, j.lay.total);
    (void)hipEventDestroy(j.ev);
    s->jobs.erase(it);
    s->cv.notify_all();
    s->done_cv.notify_all();  // waiters on this id return TVH_ERR_RANGE
    return TVH_OK;
}

int tvh_server_stats(tvh_server* s, size_t* arena_used, size_t* arena_size, size_t* queued,
                     size_t* running) {
    if (!s) return set_error(TVH_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (arena_used) *arena_used = s->arena.used();
    if (arena_size) *arena_size = s->arena.size();
    if (queued) *queued = s->pending.size();
    if (running) {
        size_t r = 0;
        for (int64_t x : s->slot_job) r += x >= 0;
        *running = r;
    }
    return TVH_OK;
}

int tvh_server_destroy(tvh_server* s) {
    if (!s) return TVH_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (uint64_t id : s->pending) {
            Job& j = s->jobs.at(id);
            j.state = TVH_JOB_FAILED;
            j.status = TVH_ERR_INVALID_ARG;
            j.err = "server destroyed before the job was admitted";
        }
        s->pending.clear();
        s->done_cv.notify_all();
        s->cv.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    for (auto st : s->slots) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    {
        // jobs still marked running never had their completion observed: fail
        // them under the lock, wake their waiters and wait until every waiter
        // has left tvh_server_wait before the condition variables go away
        std::unique_lock<std::mutex> lk(s->mu);
        for (auto& kv : s->jobs)
            if (kv.second.state == TVH_JOB_RUNNING) {
                kv.second.state = TVH_JOB_FAILED;
                kv.second.status = TVH_ERR_DEVICE;
                kv.second.err = "server destroyed while the job was running";
            }
        s->done_cv.notify_all();
        s->waiters_cv.wait(lk, [&] { return s->waiters == 0; });
    }
    (void)hipStreamSynchronize(s->qstream);
    (void)hipStreamDestroy(s->qstream);
    for (auto& kv : s->jobs)
        if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
    int rc = TVH_OK;
    hipError_t e;
    if (s->qbuf && (e = hipFree(s->qbuf)) != hipSuccess)
        rc = set_error(TVH_ERR_DEVICE, "query buffer free: %s", hipGetErrorString(e));
    if (s->root_stage && (e = hipHostFree(s->root_stage)) != hipSuccess && !rc)
        rc = set_error(TVH_ERR_DEVICE, "root staging free: %s", hipGetErrorString(e));
    // The runtime returns a large allocation to the device's free pool some
    // time after hipFree (measured: hundreds of GB reappear later, not at the
    // call).  Wait, bounded, until the free figure has grown by the arena's
    // size, so that a caller can reuse the memory as soon as this returns.
    size_t free0 = 0, total = 0, arena_bytes = s->arena.size();
    (void)hipDeviceSynchronize();
    (void)hipMemGetInfo(&free0, &total);
    if ((e = hipFree(s->base)) != hipSuccess && !rc)
        rc = set_error(TVH_ERR_DEVICE, "arena free: %s", hipGetErrorString(e));
    if (!rc) {
        const size_t want = free0 + arena_bytes - arena_bytes / 64;  // within ~1.5 %
        auto t0 = std::chrono::steady_clock::now();
        for (;;) {
            size_t fr = 0;
            if (hipMemGetInfo(&fr, &total) != hipSuccess || fr >= want) break;
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(20)) break;
            std::this_thread::sleep_for(std::chrono::milliseconds(2));
        }
    }
    delete s;
    return rc;
}

}  // extern "C"
