// C ABI of libgpufsm_hip.so (declared in include/gpufsm_hip.h).
//
// Owns the device-resident automaton handles, per-call device memory (stream-ordered
// hipMallocAsync, so concurrent calls on distinct streams never share scratch), technique
// resolution, HIP-event timing of the host forms, and the conversion of exceptions into
// status codes plus a thread-local message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpufsm_hip.h"
#include "internal.hpp"
#include "layout.hpp"

using namespace gfsm;

struct gfsm_nfa {
    int32_t n = 0, start = 0;
    int64_t nnz = 0, nne = 0;
    bool start_accepts = false;
    NfaWorklistDev wl{};
    // general layout (epsilon closure folded): uploaded for automata that are not
    // chain-structured and for any automaton of at most 512 states
    bool general_ok = false;
    std::string general_why;
    NfaGeneralDev gen{};
    bool chain_ok = false;
    NfaTileDev tile{};
    NfaChainDev chain{};
    int32_t max_chain = 0;
    std::vector<void *> blocks;
    int64_t device_bytes = 0;
    // Byte-triple lists of the chain technique's lane tier (64 MiB of headers plus the
    // lists), built on the first run that asks for them (GFSM_TUNE_CHAIN_TRIPLE): the host
    // keeps the chain layout until then, for automata the lists can apply to.
    std::mutex tri_mu;
    int tri_state = 0;  // 0: not built yet, 1: built, -1: not available (tri_why)
    std::atomic<int64_t> tri_bytes{0};
    std::string tri_why;
    NfaTriDev tri{};
    std::vector<void *> tri_blocks;
    std::unique_ptr<TileHost> tile_host;
};

struct gfsm_dfa {
    int32_t n = 0, start = 0;
    bool start_accepts = false;
    DfaDev dev{};
    bool lds_ok = false;
    std::vector<void *> blocks;
    int64_t device_bytes = 0;
};

// Scratch of the allocation-free entry point (gfsm_dfa_run_device_static): one array per
// GPU, held until the process ends. In host code `scratch` names a host-side stand-in, not
// the device array: its device address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) unsigned char scratch[GFSM_STATIC_SCRATCH_BYTES];

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

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

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Collects the arrays of one automaton and uploads them as one allocation with one copy
// (an automaton has up to 20 arrays; per-array allocations and copies dominated the plain
// per-call path). Each array starts on a 256-byte boundary.
class Packer {
  public:
    template <typename T>
    void add(const T **field, const std::vector<T> &v) {
        add(field, v.data(), v.size());
    }
    template <typename T>
    void add(const T **field, const T *src, size_t count) {
        const size_t bytes = (count ? count : 1) * sizeof(T);
        const size_t off = total_;
        total_ = align256(total_ + bytes);
        parts_.push_back({src, count * sizeof(T), off,
                          [field](char *base, size_t o) {
                              *field = reinterpret_cast<const T *>(base + o);
                          }});
    }
    // One device allocation (recorded in blocks), one host-to-device copy.
    void commit(std::vector<void *> &blocks, int64_t &device_bytes) {
        if (total_ == 0) return;
        std::vector<char> host(total_, 0);
        for (const Part &p : parts_)
            if (p.bytes) std::memcpy(host.data() + p.off, p.src, p.bytes);
        void *dev = nullptr;
        if (hipMalloc(&dev, total_) != hipSuccess)
            throw Error(GFSM_ENOMEM, "hipMalloc of " + std::to_string(total_) +
                                         " bytes for the automaton failed");
        blocks.push_back(dev);
        device_bytes += static_cast<int64_t>(total_);
        GFSM_HIP(hipMemcpy(dev, host.data(), total_, hipMemcpyHostToDevice));
        for (const Part &p : parts_) p.set(static_cast<char *>(dev), p.off);
    }

    size_t total() const { return total_; }

  private:
    struct Part {
        const void *src;
        size_t bytes, off;
        std::function<void(char *, size_t)> set;
    };
    std::vector<Part> parts_;
    size_t total_ = 0;
};

void free_blocks(std::vector<void *> &blocks) {
    for (void *p : blocks) (void)hipFree(p);
    blocks.clear();
}

// Stream-ordered device buffer released on the same stream.
struct StreamBuf {
    void *p = nullptr;
    hipStream_t s = nullptr;
    StreamBuf(size_t bytes, hipStream_t stream) : s(stream) {
        if (hipMallocAsync(&p, bytes ? bytes : 1, stream) != hipSuccess)
            throw Error(GFSM_ENOMEM, "hipMallocAsync of " + std::to_string(bytes) + " bytes failed");
    }
    ~StreamBuf() {
        if (p) (void)hipFreeAsync(p, s);
    }
    StreamBuf(const StreamBuf &) = delete;
    StreamBuf &operator=(const StreamBuf &) = delete;
};

hipStream_t host_stream() {
    thread_local hipStream_t s = nullptr;
    if (!s) GFSM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

void check_batch_args(const void *data, const int64_t *offsets, int32_t count,
                      const int32_t *flags, const int32_t *lens) {
    if (count < 0) throw Error(GFSM_EINVAL, "num_strings must be >= 0");
    if (count > 0 && (!offsets || !flags || !lens)) throw Error(GFSM_EINVAL, "null buffer");
    (void)data;
}

// The general layout's worklist (ballot compaction) runs when the layout exists and its
// bitmaps fit LDS; otherwise the worklist walks the unfolded automaton with its own
// epsilon closure.
bool use_wl2(const gfsm_nfa *h) { return h->general_ok && h->n <= 65536; }

// The bit-parallel step on the general layout (automata that are not chain-structured).
bool gather_ok(const gfsm_nfa *h) { return h->general_ok && nfa_gather_fits(h->gen); }

// GFSM_GATHER_DIAG: 0 keeps every folded edge in the gather lists, 2 forces the first and
// diagonal masks wherever they exist (tests), anything else lets the layout decide.
DiagMode gather_diag_mode() {
    const char *v = std::getenv("GFSM_GATHER_DIAG");
    if (v && v[0] == '0' && v[1] == '\0') return DiagMode::off;
    if (v && v[0] == '2' && v[1] == '\0') return DiagMode::force;
    return DiagMode::automatic;
}

int32_t resolve_nfa(const gfsm_nfa *h, int32_t technique, int32_t count) {
    switch (technique) {
        case GFSM_NFA_AUTO:
            // chain-structured automata: the sparse chain walk (its own tiers fall back to
            // the tiled pass on inputs whose active sets outgrow the lists); small general
            // automata over batches that fill the lanes: one lane per input; general
            // automata whose start loops on every byte with first masks (at least one
            // successor per 64 states per byte, so every set holds that many states and
            // never dies): the masked bit-parallel step, whose cost per byte does not grow
            // with the set; otherwise the wave-cooperative worklist
            if (h->chain_ok) return GFSM_NFA_CHAIN;
            if (h->general_ok && h->gen.nwords <= kRegMaxWords && count >= 1024)
                return GFSM_NFA_REGISTER;
            if (gather_ok(h) && h->gen.fmask) return GFSM_NFA_BITPACKED;
            return GFSM_NFA_WORKLIST;
        case GFSM_NFA_WORKLIST:
            return technique;
        case GFSM_NFA_BITPACKED:
            if (h->chain_ok || gather_ok(h)) return technique;
            throw Error(GFSM_EUNSUPPORTED,
                        "bitpacked needs a chain-structured NFA (tiled shift step) or a "
                        "general layout whose bitsets fit LDS: " +
                            (h->general_ok ? std::to_string(h->n) + " states"
                                           : h->general_why));
        case GFSM_NFA_CHAIN:
            if (!h->chain_ok)
                throw Error(GFSM_EUNSUPPORTED,
                            "chain needs a chain-structured NFA (no epsilon edges, at most "
                            "one predecessor and successor per non-start state, homogeneous "
                            "labels, chains of at most 4096 states)");
            return technique;
        case GFSM_NFA_REGISTER:
            if (!h->general_ok || h->gen.nwords > kRegMaxWords)
                throw Error(GFSM_EUNSUPPORTED,
                            "register needs at most " + std::to_string(kRegMaxWords * 64) +
                                " states with a foldable epsilon closure" +
                                (h->general_ok ? std::string() : ": " + h->general_why));
            return technique;
        default:
            throw Error(GFSM_EINVAL, "unknown NFA technique " + std::to_string(technique));
    }
}

struct Tuning {
    int32_t chunk = 0, segment = 0, flags = 0;
    int32_t max_len = 0;  // longest input, 0 = not known
};

Tuning read_tuning(const gfsm_tuning *t) {
    Tuning r;
    if (t) {
        if (t->chunk_bytes < 0 || t->segment_bytes < 0 || t->max_input_bytes < 0)
            throw Error(GFSM_EINVAL, "tuning sizes must be >= 0");
        r.chunk = t->chunk_bytes;
        r.segment = t->segment_bytes;
        r.flags = t->flags;
        r.max_len = t->max_input_bytes;
    }
    return r;
}

// A one-lane-per-input walk lasts as long as its longest input (about 25 ns per byte for a
// lane alone), the segmented walk about total / 1.3 TB/s: an input longer than
// total / 32768 bytes (and than four segments) sends the batch to the segmented walk.
constexpr int64_t kLaneWalkSkew = 32768;

int32_t resolve_dfa(const gfsm_dfa *h, int32_t technique, int32_t count, int64_t total,
                    const Tuning &tu = Tuning()) {
    switch (technique) {
        case GFSM_DFA_AUTO: {
            // Segment the batch when inputs span several segments: one lane per input
            // would serialise each of them into one long dependent chain. With 256 inputs
            // per CU or more, one lane per input already keeps the LDS busy and needs no
            // repair passes (65,536 x 16 KiB, 256 states: 1.86 TB/s against 1.69 split),
            // and for a table in L2 it keeps more lookups in flight than the segments do
            // (65,536 x 64 KiB, 4096 states: gather 261 GB/s against 186 split; at 32,768
            // inputs split still leads, 186 against 179).
            // A known longest input that is long against both the segment and the batch
            // is segmented whatever the other inputs are: one lane per input would walk it
            // alone. The count rule needs the longest input for the same reason.
            const DfaSplitPlan plan = dfa_split_plan(h->dev, total, count, tu.segment);
            const int64_t avg = count > 0 ? total / count : 0;
            if (tu.max_len > 0) {
                const int64_t bound = std::max<int64_t>(4 * plan.seg_bytes, total / kLaneWalkSkew);
                if (tu.max_len > bound) return GFSM_DFA_SPLIT;
                if (count >= 256 * static_cast<int64_t>(device_cu_count()))
                    return h->lds_ok ? GFSM_DFA_LDS : GFSM_DFA_GATHER;
            }
            if (avg >= 4 * plan.seg_bytes) return GFSM_DFA_SPLIT;
            return h->lds_ok ? GFSM_DFA_LDS : GFSM_DFA_GATHER;
        }
        case GFSM_DFA_GATHER:
        case GFSM_DFA_SPLIT:
            return technique;
        case GFSM_DFA_LDS:
            if (!h->lds_ok)
                throw Error(GFSM_EUNSUPPORTED, "lds needs a DFA of at most 320 states");
            return technique;
        default:
            throw Error(GFSM_EINVAL, "unknown DFA technique " + std::to_string(technique));
    }
}

// The handle's byte-triple lists, built and uploaded on first use (null when the automaton
// has none: not a pattern union whose start state loops on every byte, or lists over the
// caps). The build is host work plus a synchronous upload, so it is refused inside stream
// capture.
const NfaTriDev *chain_triples(gfsm_nfa *h, hipStream_t s) {
    std::lock_guard<std::mutex> lk(h->tri_mu);
    if (h->tri_state == 0) {
        if (!h->tile_host) {
            h->tri_state = -1;
            h->tri_why = "not a pattern union whose start state loops on every byte";
        } else {
            hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
            GFSM_HIP(hipStreamIsCapturing(s, &st));
            if (st != hipStreamCaptureStatusNone)
                throw Error(GFSM_EINVAL,
                            "chain byte-triple lists are built by the first run that asks for "
                            "them (GFSM_TUNE_CHAIN_TRIPLE): run it once outside stream capture");
            const TripleHost th = build_triples(*h->tile_host);
            if (th.ok) {
                Packer pk;
                pk.add(&h->tri.hdr, th.hdr);
                pk.add(&h->tri.nonempty, th.nonempty);
                pk.add(&h->tri.slot, th.slot);
                pk.add(&h->tri.pair_acc, th.pair_acc);
                int64_t bytes = 0;
                pk.commit(h->tri_blocks, bytes);
                h->tri_bytes.store(bytes);
                h->tri_state = 1;
            } else {
                h->tri_state = -1;
                h->tri_why = th.why;
            }
            h->tile_host.reset();
        }
    }
    return h->tri_state > 0 ? &h->tri : nullptr;
}

// Device scratch of one NFA run (resolved technique).
size_t nfa_workspace(const gfsm_nfa *h, int32_t tech, int32_t count, int64_t total,
                     const Tuning &tu) {
    if (count <= 0 || h->start_accepts) return 0;
    if (tech == GFSM_NFA_CHAIN)
        return nfa_chain_scratch_bytes(h->chain, count, total,
                                       nfa_chain_chunk_bytes(tu.chunk, total));
    if (tech == GFSM_NFA_BITPACKED)
        return h->chain_ok ? align256(static_cast<size_t>(count) * sizeof(int32_t)) : 0;
    if (tech == GFSM_NFA_REGISTER) return 0;
    int32_t grid = 1;
    int64_t scratch_elems = 0;
    if (use_wl2(h))
        nfa_wl2_geometry(h->gen, count, &grid, &scratch_elems);
    else
        nfa_worklist_geometry(h->wl, count, &grid, &scratch_elems);
    return align256(static_cast<size_t>(scratch_elems) * sizeof(int32_t));
}

// Enqueue one NFA run; ws == nullptr allocates the scratch stream-ordered.
void nfa_enqueue(gfsm_nfa *h, int32_t tech, const Tuning &tu, const uint8_t *d_data,
                 const int64_t *d_off, int32_t count, int64_t total, int32_t *d_flags,
                 int32_t *d_lens, void *ws, size_t ws_bytes, hipStream_t s) {
    if (count == 0) return;
    if (h->start_accepts) {
        launch_fill_result(d_flags, d_lens, count, 1, 0, s);
        return;
    }
    const size_t need = nfa_workspace(h, tech, count, total, tu);
    if (!ws && need > 0) {
        StreamBuf buf(need, s);
        nfa_enqueue(h, tech, tu, d_data, d_off, count, total, d_flags, d_lens, buf.p, need, s);
        return;
    }
    if (ws_bytes < need)
        throw Error(GFSM_EINVAL, "workspace of " + std::to_string(ws_bytes) + " bytes, " +
                                     std::to_string(need) + " needed");
    if (tech == GFSM_NFA_CHAIN) {
        const bool pairs = !(tu.flags & GFSM_TUNE_CHAIN_FIRST) && env_int("GFSM_CHAIN_FIRST", 0) != 1;
        const bool paired =
            !(tu.flags & GFSM_TUNE_CHAIN_WAVE64) && env_int("GFSM_CHAIN_WAVE64", 0) != 1;
        const bool triples =
            pairs && ((tu.flags & GFSM_TUNE_CHAIN_TRIPLE) || env_int("GFSM_CHAIN_TRIPLE", 0) == 1);
        const NfaTriDev *tri = triples ? chain_triples(h, s) : nullptr;
        launch_nfa_chain(h->chain, h->tile, d_data, d_off, count, total, d_flags, d_lens, ws,
                         nfa_chain_chunk_bytes(tu.chunk, total), pairs, paired, tri, s);
        return;
    }
    if (tech == GFSM_NFA_BITPACKED) {
        if (h->chain_ok)
            launch_nfa_bitpacked(h->tile, d_data, d_off, count, d_flags, d_lens,
                                 static_cast<int32_t *>(ws), s);
        else
            launch_nfa_gather(h->gen, d_data, d_off, count, d_flags, d_lens, s);
        return;
    }
    if (tech == GFSM_NFA_REGISTER) {
        launch_nfa_register(h->gen, d_data, d_off, count, d_flags, d_lens, s);
        return;
    }
    int32_t grid = 1;
    int64_t scratch_elems = 0;
    if (use_wl2(h)) {
        nfa_wl2_geometry(h->gen, count, &grid, &scratch_elems);
        launch_nfa_wl2(h->gen, d_data, d_off, count, d_flags, d_lens,
                       scratch_elems > 0 ? static_cast<int32_t *>(ws) : nullptr, grid, s);
        return;
    }
    nfa_worklist_geometry(h->wl, count, &grid, &scratch_elems);
    launch_nfa_worklist(h->wl, d_data, d_off, count, d_flags, d_lens,
                        scratch_elems > 0 ? static_cast<int32_t *>(ws) : nullptr, grid, s);
}

size_t dfa_workspace(const gfsm_dfa *h, int32_t tech, int32_t count, int64_t total,
                     const Tuning &tu) {
    if (count <= 0 || h->start_accepts || tech != GFSM_DFA_SPLIT) return 0;
    return dfa_split_plan(h->dev, total, count, tu.segment).scratch_bytes;
}

void dfa_enqueue(gfsm_dfa *h, int32_t tech, const Tuning &tu, const uint8_t *d_data,
                 const int64_t *d_off, int32_t count, int64_t total, int32_t *d_flags,
                 int32_t *d_lens, void *ws, size_t ws_bytes, hipStream_t s) {
    if (count == 0) return;
    if (h->start_accepts) {
        launch_fill_result(d_flags, d_lens, count, 1, 0, s);
        return;
    }
    if (tech == GFSM_DFA_SPLIT) {
        const DfaSplitPlan plan = dfa_split_plan(h->dev, total, count, tu.segment);
        if (!ws) {
            StreamBuf buf(plan.scratch_bytes, s);
            launch_dfa_split(h->dev, plan, d_data, d_off, count, total, d_flags, d_lens, buf.p, s);
            return;
        }
        if (ws_bytes < plan.scratch_bytes)
            throw Error(GFSM_EINVAL, "workspace of " + std::to_string(ws_bytes) + " bytes, " +
                                         std::to_string(plan.scratch_bytes) + " needed");
        launch_dfa_split(h->dev, plan, d_data, d_off, count, total, d_flags, d_lens, ws, s);
        return;
    }
    if (tech == GFSM_DFA_LDS)
        launch_dfa_lds(h->dev, d_data, d_off, count, d_flags, d_lens, s);
    else
        launch_dfa_gather(h->dev, d_data, d_off, count, d_flags, d_lens, s);
}

// Device address of `scratch` on the current device: looked up on the first call on each
// device (with that device current, as hipGetSymbolAddress requires) and kept.
std::mutex g_static_mu;
std::map<int, void *> g_static_ptr;

void *static_scratch() {
    int dev = 0;
    GFSM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_static_mu);
    auto it = g_static_ptr.find(dev);
    if (it != g_static_ptr.end()) return it->second;
    void *ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !ptr)
        throw Error(GFSM_EDEVICE, std::string("hipGetSymbolAddress of the static scratch failed: ") +
                                      hipGetErrorString(rc));
    g_static_ptr[dev] = ptr;
    return ptr;
}

void check_ws(const void *ws, int64_t ws_bytes) {
    if (ws_bytes < 0) throw Error(GFSM_EINVAL, "workspace_bytes must be >= 0");
    if (ws_bytes > 0 && !ws) throw Error(GFSM_EINVAL, "null workspace");
    if (reinterpret_cast<uintptr_t>(ws) & 255u)
        throw Error(GFSM_EINVAL, "workspace must be 256-byte aligned");
}

// Longest input of a host batch (the offsets are checked by run_host before any launch;
// malformed ones give a value that the check then rejects). 1 for a batch of empty inputs,
// so the value always counts as known.
int32_t longest_input(const int64_t *offsets, int32_t count) {
    int64_t m = 1;
    for (int32_t i = 0; i < count; ++i) {
        const int64_t l = offsets[i + 1] - offsets[i];
        if (l > m) m = l;
    }
    return m > INT32_MAX ? INT32_MAX : static_cast<int32_t>(m);
}

// Pinned host staging of the calling thread (grows, kept for the life of the thread).
struct HostStage {
    char *p = nullptr;
    size_t cap = 0;
    char *get(size_t bytes) {
        if (bytes > cap) {
            if (p) (void)hipHostFree(p);
            p = nullptr;
            cap = 0;
            const size_t want = bytes < (size_t(1) << 20) ? (size_t(1) << 20) : bytes;
            GFSM_HIP(hipHostMalloc(reinterpret_cast<void **>(&p), want, hipHostMallocDefault));
            cap = want;
        }
        return p;
    }
};

// Batches up to this size go through the pinned staging buffer (one copy each way);
// larger ones are copied straight from the caller's pageable memory.
constexpr size_t kStageLimit = size_t(64) << 20;

// Host form shared by NFA and DFA. One device allocation holds offsets, bytes, outputs and
// the technique's scratch; small batches cross PCIe as one staged copy each way. The launch
// is timed with HIP events, as the upstream entry points time theirs.
template <typename WsBytes, typename Enqueue>
void run_host(const uint8_t *data, const int64_t *offsets, int32_t count, int32_t *flags,
              int32_t *lens, float *kernel_ms, WsBytes &&ws_bytes, Enqueue &&enqueue) {
    if (kernel_ms) *kernel_ms = 0.0f;
    if (count == 0) return;
    const int64_t total = offsets[count] - offsets[0];
    if (offsets[0] != 0) throw Error(GFSM_EINVAL, "offsets[0] must be 0");
    for (int32_t i = 0; i < count; ++i) {
        if (offsets[i + 1] < offsets[i]) throw Error(GFSM_EINVAL, "offsets are not non-decreasing");
        if (offsets[i + 1] - offsets[i] > INT32_MAX)
            throw Error(GFSM_EINVAL, "input " + std::to_string(i) +
                                         " is longer than 2^31 - 1 bytes (match_len is int32)");
    }
    if (total > 0 && !data) throw Error(GFSM_EINVAL, "null data");
    hipStream_t s = host_stream();
    const size_t off_bytes = static_cast<size_t>(count + 1) * sizeof(int64_t);
    const size_t out_bytes = static_cast<size_t>(count) * 2 * sizeof(int32_t);
    const size_t ws = ws_bytes(total);
    // device layout: offsets | bytes | outputs | scratch
    const size_t o_off = 0, o_data = align256(off_bytes);
    const size_t o_out = o_data + align256(static_cast<size_t>(total));
    const size_t o_ws = o_out + align256(out_bytes);
    StreamBuf dev(o_ws + ws, s);
    char *base = static_cast<char *>(dev.p);
    thread_local HostStage stage;
    const bool staged = static_cast<size_t>(total) <= kStageLimit;
    // Past this point copies from the thread's pinned staging buffer may be in flight: an
    // error must not return before they finish, or the next call would overwrite the
    // buffer under them.
    try {
        if (staged) {
            char *h = stage.get(o_out > out_bytes ? o_out : out_bytes);
            std::memcpy(h + o_off, offsets, off_bytes);
            if (total > 0) std::memcpy(h + o_data, data, static_cast<size_t>(total));
            GFSM_HIP(hipMemcpyAsync(base, h, o_data + static_cast<size_t>(total),
                                    hipMemcpyHostToDevice, s));
        } else {
            GFSM_HIP(hipMemcpyAsync(base + o_off, offsets, off_bytes, hipMemcpyHostToDevice, s));
            GFSM_HIP(hipMemcpyAsync(base + o_data, data, static_cast<size_t>(total),
                                    hipMemcpyHostToDevice, s));
        }
        int32_t *d_flags = reinterpret_cast<int32_t *>(base + o_out);
        int32_t *d_lens = d_flags + count;
        thread_local hipEvent_t e0 = nullptr, e1 = nullptr;
        if (!e0) GFSM_HIP(hipEventCreate(&e0));
        if (!e1) GFSM_HIP(hipEventCreate(&e1));
        GFSM_HIP(hipEventRecord(e0, s));
        enqueue(reinterpret_cast<const uint8_t *>(base + o_data),
                reinterpret_cast<const int64_t *>(base + o_off), total, d_flags, d_lens,
                ws ? static_cast<void *>(base + o_ws) : nullptr, ws, s);
        GFSM_HIP(hipEventRecord(e1, s));
        if (staged) {
            char *h = stage.p;  // large enough: o_out >= out_bytes or sized for it above
            GFSM_HIP(hipMemcpyAsync(h, d_flags, out_bytes, hipMemcpyDeviceToHost, s));
            GFSM_HIP(hipStreamSynchronize(s));
            std::memcpy(flags, h, static_cast<size_t>(count) * sizeof(int32_t));
            std::memcpy(lens, h + static_cast<size_t>(count) * sizeof(int32_t),
                        static_cast<size_t>(count) * sizeof(int32_t));
        } else {
            GFSM_HIP(hipMemcpyAsync(flags, d_flags, static_cast<size_t>(count) * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, s));
            GFSM_HIP(hipMemcpyAsync(lens, d_lens, static_cast<size_t>(count) * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, s));
            GFSM_HIP(hipStreamSynchronize(s));
        }
        float ms = 0.0f;
        GFSM_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (kernel_ms) *kernel_ms = ms;
    } catch (...) {
        (void)hipStreamSynchronize(s);
        throw;
    }
}

// The general layout is built for automata that are not chain-structured (their techniques
// all run on it) and for any automaton of at most 512 states (the register technique; the
// fold costs little there). Chain-structured unions above that size run the chain, tiled
// and unfolded worklist walks, and skip the fold's host time in the per-call path.
bool wants_general(int32_t num_states, const TileHost &t) {
    return !t.ok || num_states <= kRegMaxWords * 64;
}

// Lays out every device array of an NFA handle in `pk` (and fills the scalar fields).
void pack_nfa(gfsm_nfa *h, int32_t num_states, const WorklistHost &w, const TileHost &t,
              const GeneralHost &g, std::vector<uint16_t> &s16, Packer &pk) {
    pk.add(&h->wl.row, w.row);
    pk.add(&h->wl.edge, w.edge);
    pk.add(&h->wl.bigidx, w.bigidx);
    pk.add(&h->wl.bigoff, w.bigoff);
    pk.add(&h->wl.erow, w.erow);
    pk.add(&h->wl.etgt, w.etgt);
    pk.add(&h->wl.accw, w.accw);
    pk.add(&h->wl.start_list, w.start_list);
    h->wl.n = num_states;
    h->wl.nwords = w.nwords;
    h->wl.k0 = static_cast<int32_t>(w.start_list.size());
    h->wl.has_eps = w.has_eps ? 1 : 0;
    h->general_ok = g.ok;
    h->general_why = g.ok ? std::string() : g.why;
    if (g.ok) {
        const WorklistHost &r = g.rows;
        pk.add(&h->gen.row, r.row);
        pk.add(&h->gen.edge, r.edge);
        pk.add(&h->gen.bigidx, r.bigidx);
        pk.add(&h->gen.bigoff, r.bigoff);
        pk.add(&h->gen.accw, r.accw);
        pk.add(&h->gen.start_list, r.start_list);
        pk.add(&h->gen.start_bits, g.start_bits);
        if (g.gather_ok) {
            pk.add(&h->gen.boff, g.boff);
            pk.add(&h->gen.bedge, g.bedge);
            h->gen.nedges = static_cast<int32_t>(g.bedge.size());
            if (!g.fmask.empty()) pk.add(&h->gen.fmask, g.fmask);
            if (g.ndiag) pk.add(&h->gen.dmask, g.dmask);
            h->gen.ndiag = g.ndiag;
            static_assert(sizeof(h->gen.doff) == sizeof(g.doff), "diagonal offsets");
            for (int k = 0; k < kMaxDiag; ++k) h->gen.doff[k] = g.doff[k];
        }
        h->gen.n = num_states;
        h->gen.nwords = r.nwords;
        h->gen.k0 = static_cast<int32_t>(r.start_list.size());
    }
    h->chain_ok = t.ok;
    if (t.ok) {
        pk.add(&h->tile.mask, t.mask);
        pk.add(&h->tile.first, t.first);
        pk.add(&h->tile.hassucc, t.hassucc);
        pk.add(&h->tile.acc, t.acc);
        for (int i = 0; i < 4; ++i) h->tile.selfmask[i] = t.selfmask[i];
        h->tile.tiles = t.tiles;
        h->tile.start_always = t.start_always ? 1 : 0;
        h->max_chain = t.max_chain;
        pk.add(&h->chain.next, t.next);
        pk.add(&h->chain.acc, t.acc);
        pk.add(&h->chain.first_off, t.first_off);
        pk.add(&h->chain.first_slot, t.first_slot);
        if (!t.pair.empty()) {
            pk.add(&h->chain.pair, t.pair);
            pk.add(&h->chain.pair_slot, t.pair_slot);
            if (static_cast<int64_t>(t.tiles) * kTileStates <= 65536) {
                // half the footprint: the register tier's L2 working set is the
                // continuation table plus these lists
                s16.assign(t.pair_slot.begin(), t.pair_slot.end());
                pk.add(&h->chain.pair_slot16, s16);
            }
        }
        pk.add(&h->chain.bflags, t.bflags);
        for (int i = 0; i < 4; ++i) {
            h->chain.selfmask[i] = t.selfmask[i];
            h->chain.first_any[i] = t.first_any[i];
            h->chain.first_acc[i] = t.first_acc[i];
        }
        h->chain.tiles = t.tiles;
        h->chain.max_chain = t.max_chain;
        h->chain.start_always = h->tile.start_always;
    }
}

}  // namespace

int gfsm::report_error(int code, const std::string &msg) { return fail(code, msg); }

extern "C" {

int gfsm_abi_version(void) { return GFSM_ABI_VERSION; }

const char *gfsm_last_error(void) { return g_last_error.c_str(); }

int gfsm_device_arch(int32_t ordinal, char *buf, int32_t buflen) {
    return guarded([&] {
        if (!buf || buflen <= 0) throw Error(GFSM_EINVAL, "null buffer");
        hipDeviceProp_t prop;
        GFSM_HIP(hipGetDeviceProperties(&prop, ordinal));
        std::strncpy(buf, prop.gcnArchName, static_cast<size_t>(buflen) - 1);
        buf[buflen - 1] = '\0';
    });
}

int gfsm_nfa_create(int32_t num_states, int32_t start_state, const uint8_t *accept,
                    const int32_t *sym_row_ptr, const int32_t *sym_targets,
                    const int32_t *sym_symbols, const int32_t *eps_row_ptr,
                    const int32_t *eps_targets, gfsm_nfa **out) {
    return guarded([&] {
        if (!out) throw Error(GFSM_EINVAL, "null out");
        *out = nullptr;
        NfaInput a{num_states, start_state, accept, sym_row_ptr, sym_targets, sym_symbols,
                   eps_row_ptr, eps_targets};
        validate_nfa(a);
        WorklistHost w = build_worklist(a);
        TileHost t = build_tiles(a);
        GeneralHost g;
        if (wants_general(num_states, t)) g = build_general(a, gather_diag_mode());
        else g.why = "not built for chain-structured automata above 512 states";
        auto *h = new gfsm_nfa();
        try {
            h->n = num_states;
            h->start = start_state;
            h->nnz = sym_row_ptr[num_states];
            h->nne = eps_row_ptr[num_states];
            h->start_accepts = w.start_accepts;
            Packer pk;
            std::vector<uint16_t> s16;
            pack_nfa(h, num_states, w, t, g, s16, pk);
            pk.commit(h->blocks, h->device_bytes);
            // the lane tier's lists can apply: keep what their build reads
            if (t.ok && t.start_always && !t.pair.empty() &&
                static_cast<int64_t>(t.tiles) * kTileStates <= 65536) {
                auto keep = std::make_unique<TileHost>();
                keep->ok = true;
                keep->start_always = true;
                keep->tiles = t.tiles;
                keep->next = std::move(t.next);
                keep->acc = std::move(t.acc);
                keep->pair = std::move(t.pair);
                keep->pair_slot = std::move(t.pair_slot);
                h->tile_host = std::move(keep);
            }
        } catch (...) {
            free_blocks(h->blocks);
            delete h;
            throw;
        }
        *out = h;
    });
}

int gfsm_nfa_destroy(gfsm_nfa *nfa) {
    return guarded([&] {
        if (!nfa) return;
        free_blocks(nfa->blocks);
        free_blocks(nfa->tri_blocks);
        delete nfa;
    });
}

int gfsm_nfa_query(const gfsm_nfa *nfa, gfsm_nfa_props *out) {
    return guarded([&] {
        if (!nfa || !out) throw Error(GFSM_EINVAL, "null argument");
        std::memset(out, 0, sizeof(*out));
        out->num_states = nfa->n;
        out->start_state = nfa->start;
        out->num_sym_edges = nfa->nnz;
        out->num_eps_edges = nfa->nne;
        out->start_accepts = nfa->start_accepts;
        out->chain_ok = nfa->chain_ok;
        out->num_tiles = nfa->chain_ok ? nfa->tile.tiles : 0;
        out->start_always = nfa->chain_ok ? nfa->tile.start_always : 0;
        out->max_chain = nfa->max_chain;
        out->general_ok = nfa->general_ok;
        // the chain lane tier's byte-triple lists count once a run has built them
        out->device_bytes = nfa->device_bytes + nfa->tri_bytes.load();
    });
}

int gfsm_nfa_analyze(int32_t num_states, int32_t start_state, const uint8_t *accept,
                     const int32_t *sym_row_ptr, const int32_t *sym_targets,
                     const int32_t *sym_symbols, const int32_t *eps_row_ptr,
                     const int32_t *eps_targets, gfsm_nfa_props *out) {
    return guarded([&] {
        if (!out) throw Error(GFSM_EINVAL, "null out");
        NfaInput a{num_states, start_state, accept, sym_row_ptr, sym_targets, sym_symbols,
                   eps_row_ptr, eps_targets};
        validate_nfa(a);
        WorklistHost w = build_worklist(a);
        TileHost t = build_tiles(a);
        GeneralHost g;
        if (wants_general(num_states, t)) g = build_general(a, gather_diag_mode());
        std::memset(out, 0, sizeof(*out));
        out->num_states = num_states;
        out->start_state = start_state;
        out->num_sym_edges = sym_row_ptr[num_states];
        out->num_eps_edges = eps_row_ptr[num_states];
        out->start_accepts = w.start_accepts;
        out->chain_ok = t.ok;
        out->num_tiles = t.ok ? t.tiles : 0;
        out->start_always = t.ok ? t.start_always : 0;
        out->max_chain = t.max_chain;
        out->general_ok = g.ok;
        gfsm_nfa tmp;
        Packer pk;
        std::vector<uint16_t> s16;
        pack_nfa(&tmp, num_states, w, t, g, s16, pk);
        const int64_t bytes = static_cast<int64_t>(pk.total());
        out->device_bytes = bytes;
    });
}

int gfsm_nfa_run_host(gfsm_nfa *nfa, int32_t technique, const uint8_t *data,
                      const int64_t *offsets, int32_t num_strings, int32_t *flags,
                      int32_t *lens, float *kernel_ms) {
    return guarded([&] {
        if (!nfa) throw Error(GFSM_EINVAL, "null handle");
        check_batch_args(data, offsets, num_strings, flags, lens);
        const int32_t tech = resolve_nfa(nfa, technique, num_strings);
        run_host(data, offsets, num_strings, flags, lens, kernel_ms,
                 [&](int64_t tot) { return nfa_workspace(nfa, tech, num_strings, tot, Tuning()); },
                 [&](const uint8_t *dd, const int64_t *doff, int64_t tot, int32_t *df,
                     int32_t *dl, void *ws, size_t wsb, hipStream_t s) {
                     nfa_enqueue(nfa, tech, Tuning(), dd, doff, num_strings, tot, df, dl, ws,
                                 wsb, s);
                 });
    });
}

int gfsm_nfa_run_device(gfsm_nfa *nfa, int32_t technique, const uint8_t *d_data,
                        const int64_t *d_offsets, int32_t num_strings, int64_t total_bytes,
                        int32_t *d_flags, int32_t *d_lens, void *stream) {
    return guarded([&] {
        if (!nfa) throw Error(GFSM_EINVAL, "null handle");
        check_batch_args(d_data, d_offsets, num_strings, d_flags, d_lens);
        if (total_bytes < 0) throw Error(GFSM_EINVAL, "total_bytes must be >= 0");
        const int32_t tech = resolve_nfa(nfa, technique, num_strings);
        nfa_enqueue(nfa, tech, Tuning(), d_data, d_offsets, num_strings, total_bytes, d_flags,
                    d_lens, nullptr, 0, static_cast<hipStream_t>(stream));
    });
}

int gfsm_nfa_workspace_size(gfsm_nfa *nfa, int32_t technique, int32_t num_strings,
                            int64_t total_bytes, const gfsm_tuning *tuning,
                            int64_t *out_bytes) {
    return guarded([&] {
        if (!nfa || !out_bytes) throw Error(GFSM_EINVAL, "null argument");
        if (num_strings < 0 || total_bytes < 0) throw Error(GFSM_EINVAL, "negative size");
        const Tuning tu = read_tuning(tuning);
        *out_bytes = static_cast<int64_t>(
            nfa_workspace(nfa, resolve_nfa(nfa, technique, num_strings), num_strings,
                          total_bytes, tu));
    });
}

int gfsm_nfa_run_device_ws(gfsm_nfa *nfa, int32_t technique, const gfsm_tuning *tuning,
                           const uint8_t *d_data, const int64_t *d_offsets,
                           int32_t num_strings, int64_t total_bytes, int32_t *d_flags,
                           int32_t *d_lens, void *d_workspace, int64_t workspace_bytes,
                           void *stream) {
    return guarded([&] {
        if (!nfa) throw Error(GFSM_EINVAL, "null handle");
        check_batch_args(d_data, d_offsets, num_strings, d_flags, d_lens);
        if (total_bytes < 0) throw Error(GFSM_EINVAL, "total_bytes must be >= 0");
        check_ws(d_workspace, workspace_bytes);
        const Tuning tu = read_tuning(tuning);
        const int32_t tech = resolve_nfa(nfa, technique, num_strings);
        const size_t need = nfa_workspace(nfa, tech, num_strings, total_bytes, tu);
        if (static_cast<size_t>(workspace_bytes) < need)
            throw Error(GFSM_EINVAL, "workspace of " + std::to_string(workspace_bytes) +
                                         " bytes, " + std::to_string(need) + " needed");
        nfa_enqueue(nfa, tech, tu, d_data, d_offsets, num_strings, total_bytes, d_flags, d_lens,
                    need > 0 ? d_workspace : nullptr, static_cast<size_t>(workspace_bytes),
                    static_cast<hipStream_t>(stream));
    });
}

int gfsm_nfa_resolve(gfsm_nfa *nfa, int32_t technique, const gfsm_tuning *tuning,
                     int32_t num_strings, int64_t total_bytes, int32_t *out_technique) {
    return guarded([&] {
        if (!nfa || !out_technique) throw Error(GFSM_EINVAL, "null argument");
        if (num_strings < 0 || total_bytes < 0) throw Error(GFSM_EINVAL, "negative size");
        (void)read_tuning(tuning);
        *out_technique = resolve_nfa(nfa, technique, num_strings);
    });
}

int gfsm_dfa_create(int32_t num_states, int32_t start_state, const uint8_t *accept,
                    const int32_t *trans, gfsm_dfa **out) {
    return guarded([&] {
        if (!out) throw Error(GFSM_EINVAL, "null out");
        *out = nullptr;
        DfaHost t = build_dfa(num_states, start_state, accept, trans);
        auto *h = new gfsm_dfa();
        try {
            h->n = num_states;
            h->start = start_state;
            h->start_accepts = t.start_accepts;
            Packer pk;
            const uint32_t *t32 = nullptr;
            const uint16_t *t16 = nullptr;
            const uint8_t *t8 = nullptr;
            if (t.esz == 4)
                pk.add(&t32, t.t32);
            else if (t.esz == 2)
                pk.add(&t16, t.t16);
            else
                pk.add(&t8, t.t8);
            pk.commit(h->blocks, h->device_bytes);
            h->dev.table = t.esz == 4   ? static_cast<const void *>(t32)
                           : t.esz == 2 ? static_cast<const void *>(t16)
                                        : static_cast<const void *>(t8);
            h->dev.n = num_states;
            h->dev.wide = t.wide ? 1 : 0;
            h->dev.esz = static_cast<int32_t>(t.esz);
            h->dev.start_entry = t.start_entry;
            h->dev.accept_from = t.accept_from;
            h->dev.col_bytes = t.col_bytes;
            h->lds_ok = !t.wide && dfa_table_bytes(h->dev) <= 160 * 1024;
        } catch (...) {
            free_blocks(h->blocks);
            delete h;
            throw;
        }
        *out = h;
    });
}

int gfsm_dfa_destroy(gfsm_dfa *dfa) {
    return guarded([&] {
        if (!dfa) return;
        free_blocks(dfa->blocks);
        delete dfa;
    });
}

int gfsm_dfa_query(const gfsm_dfa *dfa, gfsm_dfa_props *out) {
    return guarded([&] {
        if (!dfa || !out) throw Error(GFSM_EINVAL, "null argument");
        std::memset(out, 0, sizeof(*out));
        out->num_states = dfa->n;
        out->start_state = dfa->start;
        out->start_accepts = dfa->start_accepts;
        out->entry_bytes = dfa->dev.esz;
        out->lds_ok = dfa->lds_ok;
        out->device_bytes = dfa->device_bytes;
    });
}

int gfsm_dfa_run_host(gfsm_dfa *dfa, int32_t technique, const uint8_t *data,
                      const int64_t *offsets, int32_t num_strings, int32_t *flags,
                      int32_t *lens, float *kernel_ms) {
    return guarded([&] {
        if (!dfa) throw Error(GFSM_EINVAL, "null handle");
        check_batch_args(data, offsets, num_strings, flags, lens);
        const int64_t total = num_strings > 0 ? offsets[num_strings] : 0;
        Tuning tu;
        tu.max_len = longest_input(offsets, num_strings);
        const int32_t tech = resolve_dfa(dfa, technique, num_strings, total, tu);
        run_host(data, offsets, num_strings, flags, lens, kernel_ms,
                 [&](int64_t tot) { return dfa_workspace(dfa, tech, num_strings, tot, tu); },
                 [&](const uint8_t *dd, const int64_t *doff, int64_t tot, int32_t *df,
                     int32_t *dl, void *ws, size_t wsb, hipStream_t s) {
                     dfa_enqueue(dfa, tech, tu, dd, doff, num_strings, tot, df, dl, ws,
                                 wsb, s);
                 });
    });
}

int gfsm_dfa_run_device(gfsm_dfa *dfa, int32_t technique, const uint8_t *d_data,
                        const int64_t *d_offsets, int32_t num_strings, int64_t total_bytes,
                        int32_t *d_flags, int32_t *d_lens, void *stream) {
    return guarded([&] {
        if (!dfa) throw Error(GFSM_EINVAL, "null handle");
        check_batch_args(d_data, d_offsets, num_strings, d_flags, d_lens);
        if (total_bytes < 0) throw Error(GFSM_EINVAL, "total_bytes must be >= 0");
        const int32_t tech = resolve_dfa(dfa, technique, num_strings, total_bytes);
        dfa_enqueue(dfa, tech, Tuning(), d_data, d_offsets, num_strings, total_bytes, d_flags,
                    d_lens, nullptr, 0, static_cast<hipStream_t>(stream));
    });
}

int gfsm_dfa_run_device_static(gfsm_dfa *dfa, int32_t technique, const gfsm_tuning *tuning,
                               const uint8_t *d_data, const int64_t *d_offsets,
                               int32_t num_strings, int64_t total_bytes, int32_t *d_flags,
                               int32_t *d_lens, void *stream) {
    return guarded([&] {
        if (!dfa) throw Error(GFSM_EINVAL, "null handle");
        check_batch_args(d_data, d_offsets, num_strings, d_flags, d_lens);
        if (total_bytes < 0) throw Error(GFSM_EINVAL, "total_bytes must be >= 0");
        const Tuning tu = read_tuning(tuning);
        const int32_t tech = resolve_dfa(dfa, technique, num_strings, total_bytes, tu);
        const size_t need = dfa_workspace(dfa, tech, num_strings, total_bytes, tu);
        if (need > static_cast<size_t>(GFSM_STATIC_SCRATCH_BYTES))
            throw Error(GFSM_EUNSUPPORTED,
                        "the run needs " + std::to_string(need) + " bytes of scratch; the "
                        "static scratch holds " + std::to_string(GFSM_STATIC_SCRATCH_BYTES) +
                        " (use gfsm_dfa_run_device_ws with a larger workspace, or a smaller "
                        "batch)");
        void *ws = static_scratch();  // recorded on the first call on this device
        dfa_enqueue(dfa, tech, tu, d_data, d_offsets, num_strings, total_bytes, d_flags, d_lens,
                    ws, static_cast<size_t>(GFSM_STATIC_SCRATCH_BYTES),
                    static_cast<hipStream_t>(stream));
    });
}

int gfsm_dfa_workspace_size(gfsm_dfa *dfa, int32_t technique, int32_t num_strings,
                            int64_t total_bytes, const gfsm_tuning *tuning,
                            int64_t *out_bytes) {
    return guarded([&] {
        if (!dfa || !out_bytes) throw Error(GFSM_EINVAL, "null argument");
        if (num_strings < 0 || total_bytes < 0) throw Error(GFSM_EINVAL, "negative size");
        const Tuning tu = read_tuning(tuning);
        const int32_t tech = resolve_dfa(dfa, technique, num_strings, total_bytes, tu);
        *out_bytes = static_cast<int64_t>(dfa_workspace(dfa, tech, num_strings, total_bytes, tu));
    });
}

int gfsm_dfa_run_device_ws(gfsm_dfa *dfa, int32_t technique, const gfsm_tuning *tuning,
                           const uint8_t *d_data, const int64_t *d_offsets,
                           int32_t num_strings, int64_t total_bytes, int32_t *d_flags,
                           int32_t *d_lens, void *d_workspace, int64_t workspace_bytes,
                           void *stream) {
    return guarded([&] {
        if (!dfa) throw Error(GFSM_EINVAL, "null handle");
        check_batch_args(d_data, d_offsets, num_strings, d_flags, d_lens);
        if (total_bytes < 0) throw Error(GFSM_EINVAL, "total_bytes must be >= 0");
        check_ws(d_workspace, workspace_bytes);
        const Tuning tu = read_tuning(tuning);
        const int32_t tech = resolve_dfa(dfa, technique, num_strings, total_bytes, tu);
        const size_t need = dfa_workspace(dfa, tech, num_strings, total_bytes, tu);
        if (static_cast<size_t>(workspace_bytes) < need)
            throw Error(GFSM_EINVAL, "workspace of " + std::to_string(workspace_bytes) +
                                         " bytes, " + std::to_string(need) + " needed");
        dfa_enqueue(dfa, tech, tu, d_data, d_offsets, num_strings, total_bytes, d_flags, d_lens,
                    d_workspace, static_cast<size_t>(workspace_bytes),
                    static_cast<hipStream_t>(stream));
    });
}

int gfsm_dfa_resolve(gfsm_dfa *dfa, int32_t technique, const gfsm_tuning *tuning,
                     int32_t num_strings, int64_t total_bytes, int32_t *out_technique) {
    return guarded([&] {
        if (!dfa || !out_technique) throw Error(GFSM_EINVAL, "null argument");
        if (num_strings < 0 || total_bytes < 0) throw Error(GFSM_EINVAL, "negative size");
        *out_technique =
            resolve_dfa(dfa, technique, num_strings, total_bytes, read_tuning(tuning));
    });
}

}  // extern "C"
