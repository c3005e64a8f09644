// Serving loop for commit jobs: continuous batching with admission by HBM
// footprint.
//
// Submitted jobs queue in arrival order.  A worker thread
//  1. retires finished batches (roots delivered, memory returned to the pool),
//  2. admits queued jobs while their footprint (message + workspace) fits in the
//     pool, in order,
//  3. launches the admitted jobs as one batch on a free stream slot.
// Up to n_streams batches run at once, so jobs that arrive while others are on
// the device start as soon as memory and a slot free up, and their kernels
// overlap with the running batches.  Each slot owns its batch-plan buffers, so
// the loop neither allocates device memory nor synchronizes the device.
//
// Proof jobs (b64_server_submit_proof) are batched with other proof jobs: such a
// batch runs the phases of b64_batch_prove on its slot, and whenever a phase
// completes the worker does its host part (roots into the jobs' transcripts,
// challenges sampled, the next phase enqueued) before the slot moves on, so the
// rounds of proof batches interleave with everything else in flight.  Their
// phase buffers grow to the largest batch a slot has seen and are then reused.
#include <hip/hip_runtime.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "abi_util.hpp"
#include "internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t POOL_ALIGN = 256;
inline size_t pool_round(size_t x) { return (x + POOL_ALIGN - 1) & ~(POOL_ALIGN - 1); }

// First-fit allocator over one device allocation (offsets only).
class Pool {
   public:
    explicit Pool(size_t size) : size_(size) { free_[0] = size; }
    bool alloc(size_t n, size_t *off) {
        n = pool_round(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n, at = it->first + n;
                free_.erase(it);
                if (rest) free_[at] = rest;
                used_ += n;
                if (used_ > peak_) peak_ = used_;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = pool_round(n);
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t size() const { return size_; }
    size_t peak() const { return peak_; }

   private:
    size_t size_, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
};

struct Job {
    uint64_t ticket = 0;
    uint32_t n = 0;
    const uint64_t *msg = nullptr;  // caller-owned until the job is done
    b64_transcript *transcript = nullptr;  // proof jobs: caller-owned until the job is done
    std::vector<uint64_t> chal;
    std::vector<uint8_t> roots;
    size_t footprint = 0, off = 0, ws_off = 0;
    Clock::time_point t_submit, t_done;
    bool done = false;
};

constexpr uint32_t MAX_BATCH = 256;
constexpr size_t TABLE_BYTES = 4u << 20;  // per slot: plan tables for MAX_BATCH jobs
constexpr size_t ROOTS_CAP = MAX_BATCH * (B64_MAX_FOLD_ARITIES + 2);
constexpr size_t PROVE_RESERVE = 4u << 20;  // per slot: proof-phase tables (results: twice)

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done_ev = nullptr;
    b64::BatchStorage st{};
    b64_batch_plan *plan = nullptr;
    b64::ProveBuffers *pbuf = nullptr;
    b64::ProveBatch *prove = nullptr;  // a proof batch in progress
    bool ready = false;                // its enqueued phase has completed: host part due
    std::vector<std::shared_ptr<Job>> jobs;
    bool busy = false;
};

}  // namespace

struct b64_server {
    b64_server_config cfg{};
    uint8_t *d_pool = nullptr;
    std::unique_ptr<Pool> pool;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::atomic<uint64_t> n_submitted{0};  // bumped by every submit: the polling worker sees new work
    std::deque<std::shared_ptr<Job>> queue;
    std::map<uint64_t, std::shared_ptr<Job>> jobs;
    uint64_t next_ticket = 1, n_done = 0, n_batches = 0;
    bool stop = false;
    int worker_error = B64_OK;
    std::thread worker;

    ~b64_server() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        for (Slot &s : slots) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            if (s.plan) b64_batch_plan_destroy(s.plan);
            if (s.prove) b64::prove_batch_destroy(s.prove);
            if (s.pbuf) b64::prove_buffers_destroy(s.pbuf);
            if (s.done_ev) (void)hipEventDestroy(s.done_ev);
            if (s.stream) (void)hipStreamDestroy(s.stream);
            if (s.st.d_tables) (void)hipFree(s.st.d_tables);
            if (s.st.h_stage) (void)hipHostFree(s.st.h_stage);
            if (s.st.h_roots) (void)hipHostFree(s.st.h_roots);
        }
        if (d_pool) (void)hipFree(d_pool);
    }

    // A batch that cannot complete (its launch failed, or the device reported an
    // error): wait until nothing of it is in flight on its stream, return its
    // memory and free the slot without marking its jobs done; waiters get the
    // error through worker_error (called with mu held).
    void fail_slot(Slot &s, int rc) {
        (void)hipStreamSynchronize(s.stream);
        if (s.plan) {
            b64_batch_plan_destroy(s.plan);
            s.plan = nullptr;
        }
        if (s.prove) {
            b64::prove_batch_destroy(s.prove);
            s.prove = nullptr;
        }
        s.ready = false;
        for (auto &j : s.jobs)
            if (!j->done) pool->release(j->off, j->footprint);
        s.jobs.clear();
        s.busy = false;
        if (worker_error == B64_OK) worker_error = rc;
        cv_done.notify_all();
    }

    // A job delivered: its memory back to the pool (called with mu held).
    void deliver(Job &j, Clock::time_point now) {
        pool->release(j.off, j.footprint);
        j.t_done = now;
        j.done = true;
        n_done++;
    }

    // Every job of a slot delivered (called with mu held).
    void complete(Slot &s) {
        const auto now = Clock::now();
        for (auto &j : s.jobs)
            if (!j->done) deliver(*j, now);
        s.jobs.clear();
        s.busy = false;
    }

    // Retire finished batches; a proof batch whose phase completed is marked
    // ready for its host part instead (called with mu held).
    bool retire() {
        bool any = false;
        for (Slot &s : slots) {
            if (!s.busy || s.ready) continue;
            const hipError_t q = hipEventQuery(s.done_ev);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                fail_slot(s, B64_ERR_HIP);
                any = true;
                continue;
            }
            if (s.prove) {
                s.ready = true;
                any = true;
                continue;
            }
            if (s.plan) {
                const int rc = b64_batch_plan_collect(s.plan);
                if (rc != B64_OK) worker_error = rc;
                b64_batch_plan_destroy(s.plan);
                s.plan = nullptr;
            }
            complete(s);
            any = true;
        }
        if (any) cv_done.notify_all();
        return any;
    }

    // Admit in order while the pool has room (called with mu held).  Returns the
    // free slot the admitted jobs go to, or null.
    Slot *admit() {
        Slot *slot = nullptr;
        for (Slot &s : slots)
            if (!s.busy) {
                slot = &s;
                break;
            }
        if (!slot || queue.empty()) return nullptr;
        // a batch holds commit jobs or proof jobs, in arrival order
        const bool proofs = queue.front()->transcript != nullptr;
        while (!queue.empty() && slot->jobs.size() < cfg.max_batch && (queue.front()->transcript != nullptr) == proofs) {
            auto &j = queue.front();
            if (!pool->alloc(j->footprint, &j->off)) break;
            j->ws_off = j->off + pool_round(16ull << j->n);
            slot->jobs.push_back(j);
            queue.pop_front();
        }
        if (slot->jobs.empty()) return nullptr;
        slot->busy = true;
        n_batches++;
        return slot;
    }

    // Uploads the messages and enqueues the batch (called without mu: only the
    // worker touches a busy slot's jobs and buffers).
    int launch(Slot *slot) {
        if (slot->jobs[0]->transcript) {
            std::vector<b64_prove_job> descs(slot->jobs.size());
            for (size_t i = 0; i < slot->jobs.size(); i++) {
                Job &j = *slot->jobs[i];
                void *d_msg = d_pool + j.off;
                if (hipMemcpyAsync(d_msg, j.msg, 16ull << j.n, hipMemcpyHostToDevice, slot->stream) != hipSuccess)
                    return B64_ERR_HIP;
                descs[i] = {j.n, d_msg, d_pool + j.ws_off, j.transcript, j.roots.data()};
            }
            B64_TRY(b64::prove_batch_create(descs.data(), (uint32_t)descs.size(), slot->pbuf, &slot->prove));
            B64_TRY(b64::prove_batch_enqueue(slot->prove, slot->stream));
            if (hipEventRecord(slot->done_ev, slot->stream) != hipSuccess) return B64_ERR_HIP;
            return B64_OK;
        }
        std::vector<b64_job_desc> descs(slot->jobs.size());
        for (size_t i = 0; i < slot->jobs.size(); i++) {
            Job &j = *slot->jobs[i];
            void *d_msg = d_pool + j.off;
            if (hipMemcpyAsync(d_msg, j.msg, 16ull << j.n, hipMemcpyHostToDevice, slot->stream) != hipSuccess)
                return B64_ERR_HIP;
            descs[i] = {j.n, d_msg, j.chal.data(), d_pool + j.ws_off, j.roots.data()};
        }
        int rc = b64_batch_plan_build(descs.data(), (uint32_t)descs.size(), 0, &slot->st, slot->stream, &slot->plan);
        if (rc == B64_OK) rc = b64_batch_plan_enqueue(slot->plan, slot->stream);
        if (rc == B64_OK && hipEventRecord(slot->done_ev, slot->stream) != hipSuccess) rc = B64_ERR_HIP;
        return rc;
    }

    // The host part of a proof batch's completed phase, its next phase
    // enqueued, then the completed query tapes decommitted while that runs
    // (called without mu, like launch).
    int advance(Slot *slot, bool *done, std::vector<uint32_t> *finished) {
        bool more = false;
        B64_TRY(b64::prove_batch_step(slot->prove, &more));
        if (more) {
            B64_TRY(b64::prove_batch_enqueue(slot->prove, slot->stream));
            if (hipEventRecord(slot->done_ev, slot->stream) != hipSuccess) return B64_ERR_HIP;
        }
        b64::prove_batch_drain(slot->prove, finished);
        *done = !more;
        return B64_OK;
    }

    // With batches on the device, wait until one of them completes or a job is
    // submitted, polling (a timed condition-variable wait sleeps tens of
    // microseconds past its deadline, which a proof batch would pay once per
    // phase).  Gives the lock back every 200 us.  Only the worker reads or
    // writes the slots' state, so the loop reads it without the lock.
    void poll_inflight(std::unique_lock<std::mutex> &lk) {
        const uint64_t seen = n_submitted.load(std::memory_order_acquire);
        lk.unlock();
        const auto t_end = Clock::now() + std::chrono::microseconds(200);
        for (;;) {
            bool ready = false;
            for (Slot &s : slots)
                if (s.busy && !s.ready && hipEventQuery(s.done_ev) != hipErrorNotReady) ready = true;
            if (ready || n_submitted.load(std::memory_order_acquire) != seen || Clock::now() > t_end) break;
            std::this_thread::yield();
        }
        lk.lock();
    }

    void run() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            bool progressed = retire();
            // after an error no more work is admitted; waiters see worker_error
            if (worker_error != B64_OK) queue.clear();
            if (Slot *slot = admit()) {
                lk.unlock();
                const int rc = launch(slot);
                lk.lock();
                // the done event was not recorded: the slot must not be retired as finished
                if (rc != B64_OK) fail_slot(*slot, rc);
                progressed = true;
            }
            for (Slot &s : slots) {
                if (!s.busy || !s.ready) continue;
                s.ready = false;
                bool done = false;
                std::vector<uint32_t> finished;
                lk.unlock();
                const int rc = advance(&s, &done, &finished);
                lk.lock();
                // proofs completed in this step are delivered at once; the rest of
                // the batch carries on
                const auto now = Clock::now();
                for (uint32_t i : finished) deliver(*s.jobs[i], now);
                if (rc != B64_OK) {
                    fail_slot(s, rc);
                } else if (done) {
                    b64::prove_batch_destroy(s.prove);
                    s.prove = nullptr;
                    complete(s);
                }
                if (!finished.empty() || done || rc != B64_OK) cv_done.notify_all();
                progressed = true;
            }
            bool inflight = false;
            for (Slot &s : slots) inflight |= s.busy;
            if (stop && !inflight && queue.empty()) break;
            if (progressed) continue;
            if (inflight)
                poll_inflight(lk);
            else
                cv_work.wait(lk, [&] { return stop || !queue.empty(); });
        }
    }
};

extern "C" {

int b64_server_create(const b64_server_config *cfg, b64_server **out) {
    if (!cfg || !out || cfg->hbm_budget_bytes == 0 || cfg->max_batch == 0 || cfg->max_batch > MAX_BATCH ||
        cfg->n_streams == 0 || cfg->n_streams > 8)
        return B64_ERR_INVALID_ARGS;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev <= 0) return B64_ERR_NO_DEVICE;
    std::unique_ptr<b64_server> srv(new (std::nothrow) b64_server());
    if (!srv) return B64_ERR_OUT_OF_MEMORY;
    srv->cfg = *cfg;
    if (hipMalloc(&srv->d_pool, cfg->hbm_budget_bytes) != hipSuccess) return B64_ERR_OUT_OF_MEMORY;
    srv->pool.reset(new Pool(cfg->hbm_budget_bytes));
    srv->slots.resize(cfg->n_streams);
    for (Slot &s : srv->slots) {
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess) return B64_ERR_HIP;
        if (hipEventCreateWithFlags(&s.done_ev, hipEventDisableTiming) != hipSuccess) return B64_ERR_HIP;
        s.st.tables_cap = TABLE_BYTES;
        s.st.roots_cap = ROOTS_CAP;
        if (hipMalloc(&s.st.d_tables, TABLE_BYTES) != hipSuccess) return B64_ERR_OUT_OF_MEMORY;
        if (hipHostMalloc((void **)&s.st.h_stage, TABLE_BYTES, hipHostMallocCoherent) != hipSuccess)
            return B64_ERR_OUT_OF_MEMORY;
        if (hipHostMalloc((void **)&s.st.h_roots, 32 * ROOTS_CAP, hipHostMallocCoherent) != hipSuccess)
            return B64_ERR_OUT_OF_MEMORY;
        // proof batches' phase buffers: sized for a few dozen jobs' query phases
        // up front, so growth stays rare once the server runs
        s.pbuf = b64::prove_buffers_create(PROVE_RESERVE, s.stream);
        if (!s.pbuf) return B64_ERR_OUT_OF_MEMORY;
    }
    b64_server *p = srv.get();
    p->worker = std::thread([p] { p->run(); });
    *out = srv.release();
    return B64_OK;
}

static int submit(b64_server *srv, uint32_t log_msg_len, const uint64_t *msg, const uint64_t *challenges,
                  b64_transcript *transcript, uint64_t *ticket) {
    const size_t ws = b64_job_workspace_size(log_msg_len);
    if (ws == 0) return B64_ERR_INVALID_ARGS;
    auto j = std::make_shared<Job>();
    j->n = log_msg_len;
    j->msg = msg;
    j->transcript = transcript;
    if (challenges) j->chal.assign(challenges, challenges + 2ull * log_msg_len);
    j->roots.assign(32ull * b64_job_n_roots(log_msg_len), 0);
    j->footprint = pool_round(16ull << log_msg_len) + pool_round(ws);
    if (j->footprint > srv->pool->size()) return B64_ERR_ADMISSION_REJECTED;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        if (srv->worker_error != B64_OK) return srv->worker_error;
        j->ticket = srv->next_ticket++;
        j->t_submit = Clock::now();
        srv->jobs[j->ticket] = j;
        srv->queue.push_back(j);
        *ticket = j->ticket;
    }
    srv->n_submitted.fetch_add(1, std::memory_order_release);
    srv->cv_work.notify_one();
    return B64_OK;
}

int b64_server_submit(b64_server *srv, uint32_t log_msg_len, const uint64_t *msg, const uint64_t *challenges,
                      uint64_t *ticket) {
    if (!srv || !msg || !challenges || !ticket) return B64_ERR_INVALID_ARGS;
    return submit(srv, log_msg_len, msg, challenges, nullptr, ticket);
}

int b64_server_submit_proof(b64_server *srv, uint32_t log_msg_len, const uint64_t *msg, b64_transcript *transcript,
                            uint64_t *ticket) {
    if (!srv || !msg || !transcript || !ticket) return B64_ERR_INVALID_ARGS;
    return submit(srv, log_msg_len, msg, nullptr, transcript, ticket);
}

static int finish(b64_server *srv, std::shared_ptr<Job> j, uint8_t *roots_out, double *latency_s) {
    if (roots_out) memcpy(roots_out, j->roots.data(), j->roots.size());
    if (latency_s) *latency_s = std::chrono::duration<double>(j->t_done - j->t_submit).count();
    srv->jobs.erase(j->ticket);
    return B64_OK;
}

int b64_server_poll(b64_server *srv, uint64_t ticket, uint8_t *roots_out, int *done, double *latency_s) {
    if (!srv || !done) return B64_ERR_INVALID_ARGS;
    std::lock_guard<std::mutex> g(srv->mu);
    if (srv->worker_error != B64_OK) return srv->worker_error;
    auto it = srv->jobs.find(ticket);
    if (it == srv->jobs.end()) return B64_ERR_INDEX_OUT_OF_RANGE;
    *done = it->second->done ? 1 : 0;
    if (!*done) return B64_OK;
    return finish(srv, it->second, roots_out, latency_s);
}

int b64_server_wait(b64_server *srv, uint64_t ticket, uint8_t *roots_out, double *latency_s) {
    if (!srv) return B64_ERR_INVALID_ARGS;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(ticket);
    if (it == srv->jobs.end()) return B64_ERR_INDEX_OUT_OF_RANGE;
    std::shared_ptr<Job> j = it->second;
    srv->cv_done.wait(lk, [&] { return j->done || srv->worker_error != B64_OK; });
    if (srv->worker_error != B64_OK) return srv->worker_error;
    return finish(srv, j, roots_out, latency_s);
}

int b64_server_stats(b64_server *srv, uint64_t *jobs_done, uint64_t *batches, size_t *peak_bytes) {
    if (!srv) return B64_ERR_INVALID_ARGS;
    std::lock_guard<std::mutex> g(srv->mu);
    if (jobs_done) *jobs_done = srv->n_done;
    if (batches) *batches = srv->n_batches;
    if (peak_bytes) *peak_bytes = srv->pool->peak();
    return B64_OK;
}

void b64_server_destroy(b64_server *srv) { delete srv; }

}  // extern "C"
