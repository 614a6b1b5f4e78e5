// Native serving loop: continuous batching of match requests over device-resident automata.
//
// Requests (automaton index, bytes) are submitted from any host thread; the library copies
// the bytes at once, so the caller's buffer is free when gfsm_server_submit returns. One
// scheduler thread per server, started by gfsm_server_start, runs the loop:
//
//   completion  every in-flight launch slot whose completion event has fired hands its
//               verdicts (already copied back into pinned host memory) to its requests and
//               gives their bytes back to the budget; waiters are woken;
//   admission   waiting requests are admitted in arrival order while the HBM byte budget
//               (resident automata + the bytes of every admitted, unfinished request) has
//               room; a request that can never fit (longer than a slot, or than the budget
//               left by the automata) is rejected at once, one that waits longer than the
//               queue timeout is rejected as timed out;
//   batching    each free slot takes the oldest admitted request whose automaton has fewer
//               than max_inflight_per_automaton launches running, plus the admitted requests
//               of the same automaton after it, in order, while they fit the slot (bytes and
//               count; admitted requests wait in one queue per automaton), packs
//               them (uint8 bytes, int64 offsets) into the slot's pinned staging buffers
//               (bytes a caller submitted in place from pinned memory are copied to the
//               device from there) and
//               enqueues on the slot's own stream: the host-to-device copy, the resident run
//               (gfsm_*_run_device_ws, technique resolved per launch for its shape) and the
//               copy back of the verdicts, then records the slot's event.
//
// Slots have separate streams, device buffers and workspaces, so launches for different
// automata (and for one automaton, up to its in-flight cap) overlap on the device while
// new requests keep arriving. The scheduler polls the events of busy slots (hipEventQuery,
// a read of a completion signal) and otherwise sleeps on a condition variable that submit
// and stop signal; no Python and no interpreter lock is involved.
//
// No upstream counterpart: upstream has no serving path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/gpufsm_hip.h"
#include "copy_pool.hpp"
#include "internal.hpp"

using namespace gfsm;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Automaton {
    gfsm_nfa *nfa = nullptr;
    gfsm_dfa *dfa = nullptr;
    int32_t technique = 0;
    gfsm_tuning tuning{};
    int64_t resident = 0;  // device bytes of the handle
    int32_t inflight = 0;  // launches running
};

struct Req {
    int64_t id = 0;
    int32_t automaton = 0;
    std::shared_ptr<uint8_t> block;  // the library's copy (shared by a bulk submit's
                                     // requests; null for a caller buffer, GFSM_SERVE_NOCOPY)
    const uint8_t *bytes = nullptr;  // this request's bytes
    int64_t len = 0;
    double arrival = 0.0;
    bool pinned = false;  // the bytes lie in pinned host memory (a caller buffer): the
                          // launch copies them to the device directly, without staging
};

// Whether [p, p + n) lies in pinned (page-locked) host memory, judged by its two ends. A
// pageable pointer makes hipPointerGetAttributes fail; that error is cleared so that no
// later hipGetLastError reports it.
bool pinned_range(const uint8_t *p, int64_t n) {
    if (n <= 0) return false;
    for (const uint8_t *q : {p, p + (n - 1)}) {
        hipPointerAttribute_t attr{};
        const hipError_t rc = hipPointerGetAttributes(&attr, q);
        (void)hipGetLastError();
        if (rc != hipSuccess || attr.type != hipMemoryTypeHost) return false;
    }
    return true;
}

// Threads of a staging copy: GFSM_COPY_THREADS, else 8.
size_t copy_threads() {
    static const size_t t = [] {
        const int64_t v = env_int("GFSM_COPY_THREADS", 8);
        return static_cast<size_t>(v < 1 ? 1 : (v > 64 ? 64 : v));
    }();
    return t;
}

// Copies n bytes with up to copy_threads() threads, one per 2 MiB: one core's memcpy (about
// 5 GB/s into pinned memory) falls far short of the host's memory bandwidth, and bulk
// submits and large launches copy tens to hundreds of MiB while the GPU waits for them.
void copy_parallel(uint8_t *dst, const uint8_t *src, size_t n) {
    constexpr size_t kPerThread = size_t(2) << 20;
    const size_t t = std::min<size_t>(copy_threads(), n / kPerThread);
    if (t <= 1) {
        if (n) std::memcpy(dst, src, n);
        return;
    }
    std::vector<std::thread> pool;
    const size_t part = (n + t - 1) / t;
    for (size_t k = 1; k < t; ++k) {
        const size_t lo = k * part, hi = std::min(n, lo + part);
        if (lo < hi) pool.emplace_back([=] { std::memcpy(dst + lo, src + lo, hi - lo); });
    }
    std::memcpy(dst, src, std::min(n, part));
    for (auto &th : pool) th.join();
}

std::shared_ptr<uint8_t> new_block(size_t n) {
    return std::shared_ptr<uint8_t>(new uint8_t[n > 0 ? n : 1], std::default_delete<uint8_t[]>());
}

struct Record {
    gfsm_server_result r{};
    bool collected = false;  // returned by a wait (a tombstone until the front passes it)
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    // staging of one launch, pinned and on the device alike: the requests' bytes, then the
    // int64 offsets at the next 256-byte boundary; the verdicts: count flags
    // then count lengths (one copy out)
    uint8_t *h_buf = nullptr;
    int32_t *h_res = nullptr;
    uint8_t *d_buf = nullptr;
    int32_t *d_res = nullptr;
    void *d_ws = nullptr;
    int64_t ws_bytes = 0;
    bool busy = false;
    int32_t automaton = -1;
    std::vector<Req> reqs;
};

constexpr int32_t kPending = GFSM_SERVE_PENDING;

}  // namespace

struct gfsm_server {
    gfsm_server_config cfg{};
    int device = 0;
    std::vector<Automaton> automata;
    std::vector<Slot> slots;

    std::mutex mu;
    std::condition_variable cv_sched;  // submit, stop
    std::condition_variable cv_done;   // completions
    std::deque<Req> waiting;
    std::vector<std::deque<Req>> admitted;  // per automaton, in arrival order
    int64_t admitted_count = 0;
    // every request's record by id: ids are handed out in order, so record id lives at
    // records[id - first_id]; collected records at the front are dropped as they go (a
    // record never collected keeps the ones after it, as tombstones, until it is)
    std::deque<Record> records;
    int64_t first_id = 0;
    Record *record(int64_t id) {
        if (id < first_id || id - first_id >= static_cast<int64_t>(records.size()))
            return nullptr;
        Record &r = records[static_cast<size_t>(id - first_id)];
        return r.collected ? nullptr : &r;
    }
    int64_t next_id = 0;
    int64_t in_flight_bytes = 0, resident_bytes = 0;
    gfsm_server_stats stats{};
    bool started = false, stopping = false, drain = true, failed = false;
    std::string error;
    std::thread thread;
    std::unique_ptr<CopyPool> copies;  // the scheduler's staging copies

    ~gfsm_server();
    void run();
    void finish_locked(const Req &q, int32_t status, int32_t accepted, int32_t len, double t);
    void fail_all_locked(const std::string &msg);
    void launch(Slot &s);
};

namespace {

int fail(int code, const std::string &msg) { return gfsm::report_error(code, msg); }

template <typename F>
int guarded(F &&f) {
    try {
        f();
        return GFSM_OK;
    } catch (const Error &e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(GFSM_ENOMEM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(GFSM_EDEVICE, e.what());
    }
}

// A C-ABI call of this library made by the scheduler: failures become exceptions carrying
// the callee's message.
void lib_check(int rc, const char *what) {
    if (rc != GFSM_OK) throw Error(rc, std::string(what) + ": " + gfsm_last_error());
}

int64_t workspace_need(const Automaton &a, int32_t count, int64_t total,
                       const gfsm_tuning &tu) {
    int64_t need = 0;
    if (a.nfa)
        lib_check(gfsm_nfa_workspace_size(a.nfa, a.technique, count, total, &tu, &need),
                  "gfsm_nfa_workspace_size");
    else
        lib_check(gfsm_dfa_workspace_size(a.dfa, a.technique, count, total, &tu, &need),
                  "gfsm_dfa_workspace_size");
    return need;
}

void free_slot(Slot &s) {
    if (s.d_ws && s.stream) {  // stream-ordered allocation: freed on its stream
        (void)hipFreeAsync(s.d_ws, s.stream);
        (void)hipStreamSynchronize(s.stream);
        s.d_ws = nullptr;
    }
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    if (s.h_buf) (void)hipHostFree(s.h_buf);
    if (s.h_res) (void)hipHostFree(s.h_res);
    if (s.d_buf) (void)hipFree(s.d_buf);
    if (s.d_res) (void)hipFree(s.d_res);
    s = Slot();
}

}  // namespace

gfsm_server::~gfsm_server() {
    for (Slot &s : slots) free_slot(s);
}

void gfsm_server::finish_locked(const Req &q, int32_t status, int32_t accepted, int32_t len,
                                double t) {
    Record *rec = record(q.id);
    if (!rec || rec->r.status != kPending) return;
    gfsm_server_result &r = rec->r;
    r.status = status;
    r.accepted = accepted;
    r.match_len = len;
    r.done_s = t;
    switch (status) {
        case GFSM_SERVE_OK:
            ++stats.completed;
            stats.credited_bytes += q.len;
            break;
        case GFSM_SERVE_REJECTED: ++stats.rejected; break;
        case GFSM_SERVE_TIMEOUT: ++stats.timed_out; break;
        default: ++stats.errors; break;
    }
}

void gfsm_server::fail_all_locked(const std::string &msg) {
    failed = true;
    error = msg;
    const double t = now_s();
    for (const Req &q : waiting) finish_locked(q, GFSM_SERVE_ERROR, 0, 0, t);
    for (const auto &qa : admitted)
        for (const Req &q : qa) finish_locked(q, GFSM_SERVE_ERROR, 0, 0, t);
    for (Slot &s : slots)
        for (const Req &q : s.reqs) finish_locked(q, GFSM_SERVE_ERROR, 0, 0, t);
    waiting.clear();
    for (auto &qa : admitted) qa.clear();
    admitted_count = 0;
    cv_done.notify_all();
}

// Packs the slot's requests and enqueues copy in, run, copy out and the event. Called
// without the lock: the slot and its requests belong to the scheduler thread.
void gfsm_server::launch(Slot &s) {
    const Automaton &a = automata[static_cast<size_t>(s.automaton)];
    const int32_t count = static_cast<int32_t>(s.reqs.size());
    int64_t total = 0, longest = 0;
    // Requests that sit back to back in one block (a bulk submit) form one run: staged
    // through the pinned buffer, or, from a caller's pinned buffer, copied straight to the
    // device.
    struct Run {
        int64_t at, len;  // place in the launch's bytes
        const uint8_t *src;
        bool pinned;
    };
    std::vector<Run> runs;
    for (size_t i = 0; i < s.reqs.size(); ++i) {
        const Req &q = s.reqs[i];
        if (runs.empty() || q.bytes != runs.back().src + runs.back().len ||
            q.pinned != runs.back().pinned)
            runs.push_back({total, 0, q.bytes, q.pinned});
        runs.back().len += q.len;
        total += q.len;
        longest = std::max(longest, q.len);
    }
    const int64_t off_at = (total + 255) & ~int64_t(255);
    int64_t *h_off = reinterpret_cast<int64_t *>(s.h_buf + off_at);
    h_off[0] = 0;
    for (int32_t i = 0; i < count; ++i)
        h_off[i + 1] = h_off[i] + s.reqs[static_cast<size_t>(i)].len;
    const uint8_t *d_data = s.d_buf;
    const int64_t *d_off = reinterpret_cast<const int64_t *>(s.d_buf + off_at);
    gfsm_tuning tu = a.tuning;
    tu.max_input_bytes = static_cast<int32_t>(std::max<int64_t>(longest, 1));
    const int64_t need = workspace_need(a, count, total, tu);
    if (need > s.ws_bytes) {
        // rare: the workspace was sized for the largest shape at configuration; a larger
        // need grows it in stream order (no device-wide synchronisation)
        if (s.d_ws) GFSM_HIP(hipFreeAsync(s.d_ws, s.stream));
        s.d_ws = nullptr;
        s.ws_bytes = 0;
        GFSM_HIP(hipMallocAsync(&s.d_ws, static_cast<size_t>(need), s.stream));
        s.ws_bytes = need;
    }
    // Copies in, on the slot's own stream, whose earlier work (the slot's previous launch)
    // has finished: staged bytes go to the device in chunks of kStageChunk as soon as they
    // are staged, so the DMA of one chunk overlaps the staging of the next; pinned runs are
    // copied from the caller's memory; the offsets go first (a small copy after a large one
    // starts tens of microseconds after it, delaying the kernels). Launches in flight copy
    // side by side. (Ordering one launch's copies after another's, by an event wait on the
    // device, made some copy calls hold the scheduler for 5-8 ms, and ordering them on the
    // host measured no faster.)
    constexpr int64_t kStageChunk = int64_t(16) << 20;
    hipStream_t cs = s.stream;
    GFSM_HIP(hipMemcpyAsync(s.d_buf + off_at, s.h_buf + off_at,
                            static_cast<size_t>((count + 1) * 8), hipMemcpyHostToDevice, cs));
    int64_t sent = 0;  // launch bytes [0, sent) enqueued for the device
    auto send_staged = [&](int64_t end) {
        if (end > sent)
            GFSM_HIP(hipMemcpyAsync(s.d_buf + sent, s.h_buf + sent, static_cast<size_t>(end - sent),
                                    hipMemcpyHostToDevice, cs));
        sent = std::max(sent, end);
    };
    for (const Run &r : runs) {
        if (r.pinned) {
            send_staged(r.at);
            GFSM_HIP(hipMemcpyAsync(s.d_buf + r.at, r.src, static_cast<size_t>(r.len),
                                    hipMemcpyHostToDevice, cs));
            sent = r.at + r.len;
            continue;
        }
        for (int64_t lo = 0; lo < r.len; lo += kStageChunk) {
            const int64_t m = std::min(kStageChunk, r.len - lo);
            copies->copy(s.h_buf + r.at + lo, r.src + lo, static_cast<size_t>(m));
            if (r.at + lo + m - sent >= kStageChunk) send_staged(r.at + lo + m);
        }
    }
    send_staged(total);
    int32_t *d_flags = s.d_res, *d_lens = s.d_res + count;
    if (a.nfa)
        lib_check(gfsm_nfa_run_device_ws(a.nfa, a.technique, &tu, d_data, d_off, count, total,
                                         d_flags, d_lens, s.d_ws, s.ws_bytes, s.stream),
                  "gfsm_nfa_run_device_ws");
    else
        lib_check(gfsm_dfa_run_device_ws(a.dfa, a.technique, &tu, d_data, d_off, count, total,
                                         d_flags, d_lens, s.d_ws, s.ws_bytes, s.stream),
                  "gfsm_dfa_run_device_ws");
    GFSM_HIP(hipMemcpyAsync(s.h_res, s.d_res, static_cast<size_t>(count) * 8,
                            hipMemcpyDeviceToHost, s.stream));
    GFSM_HIP(hipEventRecord(s.done, s.stream));
}

void gfsm_server::run() {
    try {
        GFSM_HIP(hipSetDevice(device));
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> lk(mu);
        fail_all_locked(e.what());
        return;
    }
    std::unique_lock<std::mutex> lk(mu);
    try {
        for (;;) {
            const bool any_busy = std::any_of(slots.begin(), slots.end(),
                                              [](const Slot &s) { return s.busy; });
            if (stopping && (!drain || (!any_busy && waiting.empty() && admitted_count == 0)))
                break;
            bool progress = false;
            // completion (event queries outside the lock: they may take microseconds)
            if (any_busy) {
                std::vector<size_t> ready;
                lk.unlock();
                for (size_t i = 0; i < slots.size(); ++i) {
                    if (!slots[i].busy) continue;
                    const hipError_t q = hipEventQuery(slots[i].done);
                    if (q == hipSuccess)
                        ready.push_back(i);
                    else if (q != hipErrorNotReady)
                        throw Error(GFSM_EDEVICE, std::string("launch slot failed: ") +
                                                      hipGetErrorString(q));
                }
                lk.lock();
                const double t = now_s();
                for (size_t i : ready) {
                    Slot &s = slots[i];
                    const int32_t *flags = s.h_res;
                    const int32_t *lens = s.h_res + s.reqs.size();
                    for (size_t k = 0; k < s.reqs.size(); ++k) {
                        finish_locked(s.reqs[k], GFSM_SERVE_OK, flags[k] ? 1 : 0, lens[k], t);
                        in_flight_bytes -= s.reqs[k].len;
                    }
                    automata[static_cast<size_t>(s.automaton)].inflight -= 1;
                    s.reqs.clear();
                    s.busy = false;
                    s.automaton = -1;
                }
                if (!ready.empty()) {
                    progress = true;
                    cv_done.notify_all();
                }
            }
            // admission
            const double t = now_s();
            while (!waiting.empty()) {
                Req &q = waiting.front();
                if (q.len > cfg.max_batch_bytes ||
                    resident_bytes + q.len > cfg.budget_bytes) {
                    finish_locked(q, GFSM_SERVE_REJECTED, 0, 0, t);
                    waiting.pop_front();
                    progress = true;
                    cv_done.notify_all();
                    continue;
                }
                if (resident_bytes + in_flight_bytes + q.len <= cfg.budget_bytes) {
                    in_flight_bytes += q.len;
                    admitted[static_cast<size_t>(q.automaton)].push_back(std::move(q));
                    ++admitted_count;
                    waiting.pop_front();
                    progress = true;
                    continue;
                }
                if (cfg.queue_timeout_s > 0 && t - q.arrival > cfg.queue_timeout_s) {
                    finish_locked(q, GFSM_SERVE_TIMEOUT, 0, 0, t);
                    waiting.pop_front();
                    progress = true;
                    cv_done.notify_all();
                    continue;
                }
                break;
            }
            // batching: one launch per free slot while admitted requests can go
            for (Slot &s : slots) {
                if (s.busy || admitted_count == 0) continue;
                // the automaton whose oldest admitted request is the oldest of all (ids are
                // in submit order), among those below their in-flight cap
                int32_t ai = -1;
                for (size_t a = 0; a < admitted.size(); ++a) {
                    if (admitted[a].empty() ||
                        automata[a].inflight >= cfg.max_inflight_per_automaton)
                        continue;
                    if (ai < 0 || admitted[a].front().id <
                                      admitted[static_cast<size_t>(ai)].front().id)
                        ai = static_cast<int32_t>(a);
                }
                if (ai < 0) break;  // every automaton with requests is at its cap
                std::deque<Req> &qa = admitted[static_cast<size_t>(ai)];
                std::vector<Req> batch;
                int64_t used = 0;
                while (!qa.empty() && static_cast<int32_t>(batch.size()) < cfg.max_batch_count &&
                       (batch.empty() || used + qa.front().len <= cfg.max_batch_bytes)) {
                    used += qa.front().len;
                    batch.push_back(std::move(qa.front()));
                    qa.pop_front();
                }
                admitted_count -= static_cast<int64_t>(batch.size());
                automata[static_cast<size_t>(ai)].inflight += 1;
                stats.launches += 1;
                // the slot owns its requests from here: a failing launch fails them too
                s.reqs = std::move(batch);
                s.automaton = ai;
                s.busy = true;
                lk.unlock();
                launch(s);
                lk.lock();
                progress = true;
            }
            if (progress) continue;
            const bool busy_now = std::any_of(slots.begin(), slots.end(),
                                              [](const Slot &s) { return s.busy; });
            if (busy_now)
                cv_sched.wait_for(lk, std::chrono::microseconds(cfg.poll_us));
            else if (!waiting.empty())  // waiting for budget (timeouts still run)
                cv_sched.wait_for(lk, std::chrono::microseconds(1000));
            else if (!stopping)
                cv_sched.wait(lk);
        }
    } catch (const std::exception &e) {
        if (!lk.owns_lock()) lk.lock();
        fail_all_locked(std::string("serving loop stopped: ") + e.what());
    }
    if (!lk.owns_lock()) lk.lock();
    // a stop without drain: whatever is still queued will never run
    const double t = now_s();
    for (const Req &q : waiting) finish_locked(q, GFSM_SERVE_ERROR, 0, 0, t);
    for (const auto &qa : admitted)
        for (const Req &q : qa) finish_locked(q, GFSM_SERVE_ERROR, 0, 0, t);
    waiting.clear();
    for (auto &qa : admitted) qa.clear();
    admitted_count = 0;
    cv_done.notify_all();
    lk.unlock();
    // launches still in flight finish on the device before their slots are freed
    for (Slot &s : slots)
        if (s.busy) (void)hipEventSynchronize(s.done);
    lk.lock();
    for (Slot &s : slots) {
        if (!s.busy) continue;
        const int32_t *flags = s.h_res;
        const int32_t *lens = s.h_res + s.reqs.size();
        for (size_t k = 0; k < s.reqs.size(); ++k)
            finish_locked(s.reqs[k], failed ? GFSM_SERVE_ERROR : GFSM_SERVE_OK,
                          flags[k] ? 1 : 0, lens[k], now_s());
        s.reqs.clear();
        s.busy = false;
    }
    cv_done.notify_all();
}

extern "C" {

int gfsm_server_create(const gfsm_server_config *config, gfsm_server **out) {
    return guarded([&] {
        if (!out) throw Error(GFSM_EINVAL, "null out");
        *out = nullptr;
        gfsm_server_config c{};
        if (config) c = *config;
        if (c.slots <= 0) c.slots = 4;  // the hardware queues a process gets by default
        if (c.max_batch_count <= 0) c.max_batch_count = 4096;
        if (c.max_batch_bytes <= 0) c.max_batch_bytes = int64_t(16) << 20;
        if (c.budget_bytes <= 0) c.budget_bytes = int64_t(32) << 30;
        if (c.max_inflight_per_automaton <= 0) c.max_inflight_per_automaton = 2;
        if (c.poll_us <= 0) c.poll_us = 20;
        if (c.slots > 64) throw Error(GFSM_EINVAL, "at most 64 slots");
        if (c.max_batch_bytes > (int64_t(1) << 31) - 1)
            throw Error(GFSM_EINVAL, "max_batch_bytes must be below 2 GiB");
        auto srv = std::make_unique<gfsm_server>();
        srv->cfg = c;
        srv->copies = std::make_unique<CopyPool>(copy_threads() - 1);
        GFSM_HIP(hipGetDevice(&srv->device));
        srv->slots.resize(static_cast<size_t>(c.slots));
        const size_t cnt = static_cast<size_t>(c.max_batch_count);
        const size_t bytes = static_cast<size_t>(c.max_batch_bytes);
        for (Slot &s : srv->slots) {
            GFSM_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
            GFSM_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
            const size_t staging = bytes + 256 + (cnt + 1) * 8;
            GFSM_HIP(hipHostMalloc(reinterpret_cast<void **>(&s.h_buf), staging,
                                   hipHostMallocDefault));
            GFSM_HIP(hipHostMalloc(reinterpret_cast<void **>(&s.h_res), 2 * cnt * 4,
                                   hipHostMallocDefault));
            if (hipMalloc(reinterpret_cast<void **>(&s.d_buf), staging) != hipSuccess ||
                hipMalloc(reinterpret_cast<void **>(&s.d_res), 2 * cnt * 4) != hipSuccess)
                throw Error(GFSM_ENOMEM, "device buffers of a serving slot");
        }
        *out = srv.release();
    });
}

static int server_add(gfsm_server *s, gfsm_nfa *nfa, gfsm_dfa *dfa, int32_t technique,
                      const gfsm_tuning *tuning, int32_t *out_index) {
    return guarded([&] {
        if (!s || (!nfa && !dfa) || !out_index) throw Error(GFSM_EINVAL, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->started) throw Error(GFSM_EINVAL, "automata are added before gfsm_server_start");
        Automaton a;
        a.nfa = nfa;
        a.dfa = dfa;
        a.technique = technique;
        if (tuning) a.tuning = *tuning;
        if (nfa) {
            gfsm_nfa_props p{};
            lib_check(gfsm_nfa_query(nfa, &p), "gfsm_nfa_query");
            a.resident = p.device_bytes;
        } else {
            gfsm_dfa_props p{};
            lib_check(gfsm_dfa_query(dfa, &p), "gfsm_dfa_query");
            a.resident = p.device_bytes;
        }
        if (s->resident_bytes + a.resident > s->cfg.budget_bytes)
            throw Error(GFSM_EINVAL, "the resident automata exceed the HBM budget");
        // size every slot's workspace for the largest launch shapes of this automaton
        // (many short inputs, one long input); a launch needing more grows it
        gfsm_tuning tu = a.tuning;
        int64_t need = 0;
        for (const auto &shape : {std::make_pair(s->cfg.max_batch_count, s->cfg.max_batch_bytes),
                                  std::make_pair(int32_t(1), s->cfg.max_batch_bytes)}) {
            tu.max_input_bytes = static_cast<int32_t>(
                shape.first == 1 ? shape.second
                                 : std::max<int64_t>(1, shape.second / shape.first));
            need = std::max(need, workspace_need(a, shape.first, shape.second, tu));
        }
        for (Slot &sl : s->slots) {
            if (need <= sl.ws_bytes) continue;
            if (sl.d_ws) GFSM_HIP(hipFreeAsync(sl.d_ws, sl.stream));
            sl.d_ws = nullptr;
            sl.ws_bytes = 0;
            if (hipMallocAsync(&sl.d_ws, static_cast<size_t>(need), sl.stream) != hipSuccess)
                throw Error(GFSM_ENOMEM, "workspace of a serving slot");
            sl.ws_bytes = need;
            GFSM_HIP(hipStreamSynchronize(sl.stream));
        }
        s->resident_bytes += a.resident;
        s->automata.push_back(a);
        *out_index = static_cast<int32_t>(s->automata.size() - 1);
    });
}

int gfsm_server_add_nfa(gfsm_server *s, gfsm_nfa *nfa, int32_t technique,
                        const gfsm_tuning *tuning, int32_t *out_index) {
    return nfa ? server_add(s, nfa, nullptr, technique, tuning, out_index)
               : fail(GFSM_EINVAL, "null handle");
}

int gfsm_server_add_dfa(gfsm_server *s, gfsm_dfa *dfa, int32_t technique,
                        const gfsm_tuning *tuning, int32_t *out_index) {
    return dfa ? server_add(s, nullptr, dfa, technique, tuning, out_index)
               : fail(GFSM_EINVAL, "null handle");
}

int gfsm_server_start(gfsm_server *s) {
    return guarded([&] {
        if (!s) throw Error(GFSM_EINVAL, "null server");
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->started) throw Error(GFSM_EINVAL, "already started");
        if (s->automata.empty()) throw Error(GFSM_EINVAL, "no automata");
        s->admitted.resize(s->automata.size());
        s->started = true;
        s->thread = std::thread([s] { s->run(); });
    });
}

int gfsm_server_submit(gfsm_server *s, int32_t automaton, const uint8_t *data, int64_t len,
                       int64_t *out_id) {
    return guarded([&] {
        if (!s || !out_id || (len > 0 && !data)) throw Error(GFSM_EINVAL, "null argument");
        if (len < 0 || len > (int64_t(1) << 31) - 1)
            throw Error(GFSM_EINVAL, "request length must be in [0, 2^31 - 1]");
        Req q;
        q.automaton = automaton;
        q.len = len;
        q.block = new_block(static_cast<size_t>(len));
        q.bytes = q.block.get();
        if (len) std::memcpy(q.block.get(), data, static_cast<size_t>(len));
        q.arrival = now_s();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (automaton < 0 || automaton >= static_cast<int32_t>(s->automata.size()))
                throw Error(GFSM_EINVAL, "unknown automaton index " + std::to_string(automaton));
            if (!s->started || s->stopping) throw Error(GFSM_EINVAL, "the server is not running");
            if (s->failed) throw Error(GFSM_EDEVICE, s->error);
            q.id = s->next_id++;
            Record rec;
            rec.r.status = kPending;
            rec.r.automaton = automaton;
            rec.r.arrival_s = q.arrival;
            rec.r.bytes = len;
            s->records.push_back(rec);
            ++s->stats.submitted;
            *out_id = q.id;
            s->waiting.push_back(std::move(q));
        }
        s->cv_sched.notify_one();
    });
}

int gfsm_server_submit_many(gfsm_server *s, int32_t automaton, const uint8_t *data,
                            const int64_t *offsets, int32_t n, int32_t flags,
                            int64_t *out_ids) {
    return guarded([&] {
        if (!s || n < 0 || (n > 0 && (!offsets || !out_ids)))
            throw Error(GFSM_EINVAL, "null argument");
        if (n == 0) return;
        const int64_t first = offsets[0], total = offsets[n] - offsets[0];
        if (first < 0) throw Error(GFSM_EINVAL, "offsets must start at 0 or above");
        for (int32_t i = 0; i < n; ++i) {
            const int64_t len = offsets[i + 1] - offsets[i];
            if (len < 0 || len > (int64_t(1) << 31) - 1)
                throw Error(GFSM_EINVAL, "offsets must not decrease and requests must be at "
                                         "most 2^31 - 1 bytes");
        }
        if (total > 0 && !data) throw Error(GFSM_EINVAL, "null data");
        if (flags & ~GFSM_SERVE_NOCOPY) throw Error(GFSM_EINVAL, "unknown submit flags");
        // one copy of the whole run, shared by the requests; or none, the caller's buffer
        // staying valid until every one of them has finished
        std::shared_ptr<uint8_t> block;
        const uint8_t *src = data + first;
        bool pinned = false;
        if (!(flags & GFSM_SERVE_NOCOPY)) {
            block = new_block(static_cast<size_t>(total));
            copy_parallel(block.get(), src, static_cast<size_t>(total));
            src = block.get();
        } else {
            pinned = pinned_range(src, total);
        }
        const double t = now_s();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (automaton < 0 || automaton >= static_cast<int32_t>(s->automata.size()))
                throw Error(GFSM_EINVAL, "unknown automaton index " + std::to_string(automaton));
            if (!s->started || s->stopping) throw Error(GFSM_EINVAL, "the server is not running");
            if (s->failed) throw Error(GFSM_EDEVICE, s->error);
            for (int32_t i = 0; i < n; ++i) {
                Req q;
                q.automaton = automaton;
                q.len = offsets[i + 1] - offsets[i];
                q.block = block;
                q.bytes = src + (offsets[i] - first);
                q.arrival = t;
                q.pinned = pinned;
                q.id = s->next_id++;
                Record rec;
                rec.r.status = kPending;
                rec.r.automaton = automaton;
                rec.r.arrival_s = t;
                rec.r.bytes = q.len;
                s->records.push_back(rec);
                out_ids[i] = q.id;
                s->waiting.push_back(std::move(q));
            }
            s->stats.submitted += n;
        }
        s->cv_sched.notify_one();
    });
}

int gfsm_server_wait(gfsm_server *s, const int64_t *ids, int64_t n, double timeout_s,
                     gfsm_server_result *out) {
    return guarded([&] {
        if (!s || (n > 0 && (!ids || !out)) || n < 0) throw Error(GFSM_EINVAL, "null argument");
        std::unique_lock<std::mutex> lk(s->mu);
        for (int64_t k = 0; k < n; ++k)
            if (!s->record(ids[k]))
                throw Error(GFSM_EINVAL, "unknown or already collected request id " +
                                             std::to_string(ids[k]));
        // an id collected meanwhile by another wait counts as done (reported below). The
        // scan resumes at the first id not yet seen done (ids finish roughly in order), so a
        // wait on many ids costs about one pass over them, not one pass per wake-up.
        int64_t first = 0;
        auto all_done = [&] {
            for (; first < n; ++first) {
                const Record *r = s->record(ids[first]);
                if (r && r->r.status == kPending) return false;
            }
            return true;
        };
        if (timeout_s < 0)
            s->cv_done.wait(lk, all_done);
        else if (timeout_s > 0)
            s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), all_done);
        bool lost = false;
        for (int64_t k = 0; k < n; ++k) {
            Record *r = s->record(ids[k]);
            if (!r) {
                std::memset(&out[k], 0, sizeof(out[k]));
                out[k].status = GFSM_SERVE_ERROR;
                lost = true;
                continue;
            }
            out[k] = r->r;
            if (r->r.status != kPending) r->collected = true;
        }
        // drop the collected records at the front (ids below first_id are unknown from now)
        while (!s->records.empty() && s->records.front().collected) {
            s->records.pop_front();
            ++s->first_id;
        }
        if (lost) throw Error(GFSM_EINVAL, "a request was collected by another wait");
    });
}

int gfsm_server_query(gfsm_server *s, gfsm_server_stats *out) {
    return guarded([&] {
        if (!s || !out) throw Error(GFSM_EINVAL, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        *out = s->stats;
        out->waiting = static_cast<int64_t>(s->waiting.size());
        out->admitted = s->admitted_count;
        out->in_flight_bytes = s->in_flight_bytes;
        out->resident_bytes = s->resident_bytes;
        out->failed = s->failed ? 1 : 0;
    });
}

int gfsm_server_error(gfsm_server *s, char *buf, int32_t buflen) {
    return guarded([&] {
        if (!s || !buf || buflen <= 0) throw Error(GFSM_EINVAL, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        const size_t k = std::min(s->error.size(), static_cast<size_t>(buflen - 1));
        std::memcpy(buf, s->error.data(), k);
        buf[k] = '\0';
    });
}

int gfsm_server_stop(gfsm_server *s, int32_t drain) {
    return guarded([&] {
        if (!s) throw Error(GFSM_EINVAL, "null server");
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (!s->started || s->stopping) return;
            s->stopping = true;
            s->drain = drain != 0;
        }
        s->cv_sched.notify_all();
        if (s->thread.joinable()) s->thread.join();
    });
}

int gfsm_server_destroy(gfsm_server *s) {
    return guarded([&] {
        if (!s) return;
        (void)gfsm_server_stop(s, 0);
        delete s;
    });
}

}  // extern "C"
