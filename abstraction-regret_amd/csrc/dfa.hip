// DFA walk kernels for gfx950.
//
// Semantics: `src/gpufsm/reference.py:68-81` (simulate_dfa), latch-first-match. The
// upstream CUDA kernel is `native/dfa.cu:9-29`: one thread per string, int32 symbols, a
// separate int8 accept array, so two dependent loads and a branch per byte.
//
// Table layout (built by layout.cpp): states renumbered so that accepting states come last,
// table stored by byte (column c = successors of all states on byte c), each entry the
// successor's byte offset in a column. One step is
//     e = *(T *)(table + byte * col_bytes + e)       (v_mad_u32_u24 + one load)
// and "an accepting state was entered" is "e >= accept_from", tracked with one v_max per
// byte and tested once per block of bytes; a block with a hit is replayed byte by byte to
// find the exact position (rare by construction: it happens once per accepting input).
//
// Techniques
//   gather  one lane per input, table in global memory (uint16 entries up to 32767 states:
//           a 4096-state table is 2 MiB and stays in each XCD's 4 MB L2);
//   lds     the same walk with the table staged in LDS (tables up to 160 KiB: 320 states);
//   split   in-input parallel walk over fixed segments of the batch (see below).
#include "internal.hpp"

#include <climits>

namespace gfsm {

// ---------------------------------------------------------------- step primitives -------

template <typename T>
__device__ __forceinline__ uint32_t dstep(const char *tab, uint32_t col, uint32_t e,
                                          uint32_t b) {
    if constexpr (sizeof(T) == 2)
        return *reinterpret_cast<const T *>(tab + (__umul24(b, col) + e));
    else
        return *reinterpret_cast<const T *>(tab + (b * col + e));
}

// 16 bytes of a vector, tracking the largest entry reached.
template <typename T>
__device__ __forceinline__ uint32_t step16(const char *tab, uint32_t col, uint32_t e,
                                           const uint4 &v, uint32_t &mx) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        e = dstep<T>(tab, col, e, (w[k >> 2] >> ((k & 3) * 8)) & 0xffu);
        mx = e > mx ? e : mx;
    }
    return e;
}

// Replay 16 bytes from e; index of the first step entering an accepting state (0..15).
template <typename T>
__device__ __forceinline__ int32_t locate16(const char *tab, uint32_t col, uint32_t e,
                                            const uint4 &v, uint32_t acc) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    for (int k = 0; k < 16; ++k) {
        e = dstep<T>(tab, col, e, (w[k >> 2] >> ((k & 3) * 8)) & 0xffu);
        if (e >= acc) return k;
    }
    return 15;  // unreachable when the block's maximum was >= acc
}

__device__ __forceinline__ uint4 ld16(const uint8_t *p) {
    return *reinterpret_cast<const uint4 *>(p);
}

// Exact walk of [a, b) from e: absolute position of the first byte after which an
// accepting state is active, or -1 (then e is the state after b). Never reads outside.
template <typename T>
__device__ __forceinline__ int64_t walk_exact(const char *tab, uint32_t col, uint32_t acc,
                                              uint32_t &e, const uint8_t *__restrict__ data,
                                              int64_t a, int64_t b) {
    int64_t q = a;
    for (; q < b && (reinterpret_cast<uintptr_t>(data + q) & 15); ++q) {
        e = dstep<T>(tab, col, e, data[q]);
        if (e >= acc) return q;
    }
    for (; q + 16 <= b; q += 16) {
        const uint4 v = ld16(data + q);
        const uint32_t e0 = e;
        uint32_t mx = 0;
        e = step16<T>(tab, col, e, v, mx);
        if (mx >= acc) return q + locate16<T>(tab, col, e0, v, acc);
    }
    for (; q < b; ++q) {
        e = dstep<T>(tab, col, e, data[q]);
        if (e >= acc) return q;
    }
    return -1;
}

__global__ void fill_result_kernel(int32_t *flags, int32_t *lens, int32_t count, int32_t flag,
                                   int32_t len) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        flags[i] = flag;
        lens[i] = len;
    }
}

template <typename T, bool kLds>
__device__ __forceinline__ const char *stage_table(const T *gtab, uint32_t bytes,
                                                   uint4 *lds_raw) {
    if constexpr (kLds) {
        const uint4 *src = reinterpret_cast<const uint4 *>(gtab);
        const uint32_t vecs = bytes / 16;  // tables are n * 256 entries: a multiple of 16 B
        for (uint32_t k = threadIdx.x; k < vecs; k += blockDim.x) lds_raw[k] = src[k];
        __syncthreads();
        return reinterpret_cast<const char *>(lds_raw);
    } else {
        (void)bytes;
        (void)lds_raw;
        return reinterpret_cast<const char *>(gtab);
    }
}

// ---------------------------------------------------------------- one lane per input ----

template <typename T, bool kLds>
__global__ __launch_bounds__(256) void dfa_walk_kernel(const T *__restrict__ gtab,
                                                       uint32_t table_bytes, uint32_t col,
                                                       uint32_t start, uint32_t acc,
                                                       const uint8_t *__restrict__ data,
                                                       const int64_t *__restrict__ off,
                                                       int32_t count, int32_t *flags,
                                                       int32_t *lens) {
    extern __shared__ uint4 lds_raw[];
    const char *tab = stage_table<T, kLds>(gtab, table_bytes, lds_raw);
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += gridDim.x * blockDim.x) {
        const int64_t ib = off[i];
        uint32_t e = start;
        const int64_t hit = walk_exact<T>(tab, col, acc, e, data, ib, off[i + 1]);
        flags[i] = hit >= 0;
        lens[i] = hit >= 0 ? static_cast<int32_t>(hit - ib + 1) : 0;
    }
}

// ---------------------------------------------------------------- split walk -------------
//
// Exact in-input parallelism (the chunked DFA run of Mytkowicz, Musuvathi and Schulte,
// "Data-parallel finite-state machines", ASPLOS 2014, specialised to latch-first-match).
// The concatenated batch is cut into fixed segments of C bytes, one lane per segment.
//
//   pass 1 (walk)     Inputs that start inside the segment are walked exactly from the
//                     start state; their accepts are final. A segment that begins inside an
//                     input (its "head") is walked from the start state as a guess, storing
//                     every R bytes a checkpoint: the state reached, with bit 0 set when an
//                     accepting state was entered in that interval. A head stops early once
//                     its input is known to have accepted.
//   pass 2 (fix)      Each head is re-walked from its probable true entry (the previous
//                     segment's exit for the same input: exact if the input started there,
//                     else that segment's head, which has merged with the true path unless
//                     it never met it) until its state equals a checkpoint. From there both
//                     paths coincide and the checkpoints give the rest of the verdict.
//                     Random DFAs merge after about num_states bytes.
//   pass 3 (event)    Per head: did pass 2 assume the entry that the previous segment
//                     actually produced? A wrong assumption or an accept is an event; the
//                     earliest event of each input is kept with a 64-bit atomicMin.
//   pass 4 (finish)   Per input: an accept event gives the verdict; a wrong assumption
//                     (rare) is resolved by walking on from there, again merging with the
//                     checkpoints. The verdict is therefore always the sequential one.

struct SplitArgs {
    const uint8_t *data;
    const int64_t *off;
    int32_t count;
    int32_t rshift;  // log2 R
    int64_t total, seg, nck, segments;
    uint32_t start, acc, col, table_bytes;
    int32_t *flags, *lens, *done;
    unsigned long long *first_event;
    uint32_t *head_exit, *tail_exit, *fixed_entry, *fixed_exit;
    int32_t *fixed_acc;
    void *ckpt;  // T [segments * nck]
};

constexpr unsigned long long kNoEvent = ~0ull;

__device__ __forceinline__ int32_t split_find_input(const int64_t *__restrict__ off,
                                                    int32_t count, int64_t pos) {
    int32_t lo = 0, hi = count;  // off[lo] <= pos < off[count]
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (off[mid] <= pos) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void split_init_kernel(int32_t *flags, int32_t *lens, int32_t *done,
                                  unsigned long long *first_event, int32_t count) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        flags[i] = 0;
        lens[i] = 0;
        done[i] = 0;
        first_event[i] = kNoEvent;
    }
}

template <typename T, bool kLds>
__global__ __launch_bounds__(1024) void dfa_split_walk_kernel(const T *__restrict__ gtab,
                                                              SplitArgs a) {
    extern __shared__ uint4 lds_raw[];
    const char *tab = stage_table<T, kLds>(gtab, a.table_bytes, lds_raw);
    const uint32_t col = a.col, acc = a.acc;
    const int64_t rmask = (1ll << a.rshift) - 1;
    T *ckpt = static_cast<T *>(a.ckpt);
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < a.segments;
         g += stride) {
        const int64_t s0 = g * a.seg;
        const int64_t s1 = (s0 + a.seg) < a.total ? (s0 + a.seg) : a.total;
        int32_t i = split_find_input(a.off, a.count, s0);
        int64_t q = s0;
        // ---- head: speculative walk with checkpoints (the segment starts inside input i)
        if (a.off[i] < s0) {
            const int64_t hend = a.off[i + 1] < s1 ? a.off[i + 1] : s1;
            T *ck = ckpt + g * a.nck;
            uint32_t e = a.start, mx = 0;
            bool abandoned = false;
            // 64-byte chunks (segment starts are 64-aligned), next chunk prefetched
            uint4 v0 = make_uint4(0, 0, 0, 0), v1 = v0, v2 = v0, v3 = v0;
            if (q + 64 <= hend) {
                v0 = ld16(a.data + q);
                v1 = ld16(a.data + q + 16);
                v2 = ld16(a.data + q + 32);
                v3 = ld16(a.data + q + 48);
            }
            while (q + 64 <= hend) {
                uint4 n0 = v0, n1 = v1, n2 = v2, n3 = v3;
                if (q + 128 <= hend) {
                    n0 = ld16(a.data + q + 64);
                    n1 = ld16(a.data + q + 80);
                    n2 = ld16(a.data + q + 96);
                    n3 = ld16(a.data + q + 112);
                }
                e = step16<T>(tab, col, e, v0, mx);
                e = step16<T>(tab, col, e, v1, mx);
                e = step16<T>(tab, col, e, v2, mx);
                e = step16<T>(tab, col, e, v3, mx);
                q += 64;
                const int64_t r = q - s0;
                if ((r & rmask) == 0) {
                    ck[(r >> a.rshift) - 1] = static_cast<T>(e | (mx >= acc ? 1u : 0u));
                    mx = 0;
                }
                if ((r & 255) == 0 && __atomic_load_n(&a.done[i], __ATOMIC_RELAXED)) {
                    abandoned = true;
                    break;
                }
                v0 = n0;
                v1 = n1;
                v2 = n2;
                v3 = n3;
            }
            if (!abandoned) {
                for (; q < hend; ++q) {
                    e = dstep<T>(tab, col, e, a.data[q]);
                    mx = e > mx ? e : mx;
                    const int64_t r = q + 1 - s0;
                    if ((r & rmask) == 0) {
                        ck[(r >> a.rshift) - 1] = static_cast<T>(e | (mx >= acc ? 1u : 0u));
                        mx = 0;
                    }
                }
                const int64_t hl = hend - s0;
                if (hl & rmask) ck[hl >> a.rshift] = static_cast<T>(e | (mx >= acc ? 1u : 0u));
                a.head_exit[g] = e;
            }
            q = hend;
            ++i;
        }
        // ---- inputs that start in this segment: exact
        while (q < s1) {
            while (a.off[i + 1] <= q) ++i;  // empty inputs that start at q
            const int64_t ie = a.off[i + 1];
            const int64_t wend = ie < s1 ? ie : s1;
            uint32_t e = a.start;
            const int64_t hit = walk_exact<T>(tab, col, acc, e, a.data, q, wend);
            if (hit >= 0) {
                a.flags[i] = 1;
                a.lens[i] = static_cast<int32_t>(hit - a.off[i] + 1);
                __atomic_store_n(&a.done[i], 1, __ATOMIC_RELAXED);
            } else if (ie > s1) {
                a.tail_exit[g] = e;
            }
            q = wend;
            ++i;
        }
    }
}

// Re-walk the head [s0, hend) from `entry` until it merges with the checkpoints `ck`.
// acc_out: first accept offset relative to s0 (or -1); exit_out: state at hend.
template <typename T>
__device__ __forceinline__ void split_merge_walk(const char *tab, const SplitArgs &a,
                                                 int64_t s0, int64_t hend, uint32_t entry,
                                                 const T *ck, uint32_t spec_exit,
                                                 int32_t &acc_out, uint32_t &exit_out) {
    const uint32_t col = a.col, acc = a.acc;
    const int64_t rmask = (1ll << a.rshift) - 1;
    const uint8_t *data = a.data;
    uint32_t e = entry;
    int64_t q = s0;
    int64_t merged = -1;  // index of the checkpoint where the paths met
    acc_out = -1;
    while (q + 16 <= hend) {
        const uint4 v = ld16(data + q);
        const uint32_t e0 = e;
        uint32_t mx = 0;
        e = step16<T>(tab, col, e, v, mx);
        if (mx >= acc) {
            acc_out = static_cast<int32_t>(q - s0 + locate16<T>(tab, col, e0, v, acc));
            exit_out = e;
            return;
        }
        q += 16;
        const int64_t r = q - s0;
        if ((r & rmask) == 0 && (static_cast<uint32_t>(ck[(r >> a.rshift) - 1]) & ~1u) == e) {
            merged = (r >> a.rshift) - 1;
            break;
        }
    }
    if (merged < 0) {
        for (; q < hend; ++q) {
            e = dstep<T>(tab, col, e, data[q]);
            if (e >= acc) {
                acc_out = static_cast<int32_t>(q - s0);
                exit_out = e;
                return;
            }
        }
        exit_out = e;
        const int64_t hl = hend - s0;
        if (!(hl & rmask) || (static_cast<uint32_t>(ck[hl >> a.rshift]) & ~1u) != e) return;
        merged = hl >> a.rshift;  // met at the final partial interval: nothing after it
    }
    // paths coincide after checkpoint `merged`: take the first flagged interval after it
    const int64_t nk = ((hend - s0) + rmask) >> a.rshift;
    int64_t k = merged + 1;
    while (k < nk && !(static_cast<uint32_t>(ck[k]) & 1u)) ++k;
    if (k < nk) {
        uint32_t c = static_cast<uint32_t>(ck[k - 1]) & ~1u;
        const int64_t lo = s0 + (k << a.rshift);
        const int64_t hi = (lo + rmask + 1) < hend ? (lo + rmask + 1) : hend;
        const int64_t hit = walk_exact<T>(tab, col, acc, c, data, lo, hi);
        acc_out = static_cast<int32_t>(hit - s0);  // the flag guarantees a hit
    }
    exit_out = spec_exit;
}

// repair = 0: pass 2 for every head. repair = 1: again, only for heads whose assumed entry
// differs from the previous segment's pass-2 exit (the previous head did not merge), now
// starting from that exit. A head being repaired can race with the repair of its
// predecessor; the event pass re-checks every head afterwards, so this only affects speed.
template <typename T, bool kLds>
__global__ __launch_bounds__(1024) void dfa_split_fix_kernel(const T *__restrict__ gtab,
                                                             SplitArgs a, int32_t repair) {
    extern __shared__ uint4 lds_raw[];
    const char *tab = stage_table<T, kLds>(gtab, a.table_bytes, lds_raw);
    const T *ckpt = static_cast<const T *>(a.ckpt);
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t g = 1 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         g < a.segments; g += stride) {
        const int64_t s0 = g * a.seg;
        const int32_t i = split_find_input(a.off, a.count, s0);
        if (a.off[i] >= s0 || a.done[i]) continue;
        const int64_t s1 = (s0 + a.seg) < a.total ? (s0 + a.seg) : a.total;
        const int64_t hend = a.off[i + 1] < s1 ? a.off[i + 1] : s1;
        const bool started_before = a.off[i] >= (g - 1) * a.seg;
        uint32_t entry;
        if (repair) {
            entry = started_before ? a.tail_exit[g - 1]
                                   : __atomic_load_n(&a.fixed_exit[g - 1], __ATOMIC_RELAXED);
            if (entry == a.fixed_entry[g]) continue;
        } else {
            entry = started_before ? a.tail_exit[g - 1] : a.head_exit[g - 1];
        }
        int32_t acc_at;
        uint32_t ex;
        split_merge_walk<T>(tab, a, s0, hend, entry, ckpt + g * a.nck, a.head_exit[g], acc_at,
                            ex);
        a.fixed_entry[g] = entry;
        a.fixed_acc[g] = acc_at;
        __atomic_store_n(&a.fixed_exit[g], ex, __ATOMIC_RELAXED);
    }
}

__global__ __launch_bounds__(256) void dfa_split_event_kernel(SplitArgs a) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t g = 1 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         g < a.segments; g += stride) {
        const int64_t s0 = g * a.seg;
        const int32_t i = split_find_input(a.off, a.count, s0);
        if (a.off[i] >= s0 || a.done[i]) continue;
        const uint32_t prev =
            a.off[i] >= (g - 1) * a.seg ? a.tail_exit[g - 1] : a.fixed_exit[g - 1];
        unsigned long long key = kNoEvent;
        if (a.fixed_entry[g] != prev) key = 2ull * static_cast<unsigned long long>(g);
        else if (a.fixed_acc[g] >= 0) key = 2ull * static_cast<unsigned long long>(g) + 1ull;
        if (key != kNoEvent) atomicMin(&a.first_event[i], key);
    }
}

template <typename T, bool kLds>
__global__ __launch_bounds__(256) void dfa_split_finish_kernel(const T *__restrict__ gtab,
                                                               SplitArgs a) {
    extern __shared__ uint4 lds_raw[];
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const char *tab = reinterpret_cast<const char *>(gtab);
    if constexpr (kLds) {
        // stage the table only when some input of the block still needs a re-walk (the
        // condition is block-uniform, so every thread reaches the barrier inside)
        const bool need = i < a.count && !a.done[i] && a.first_event[i] != kNoEvent &&
                          (a.first_event[i] & 1ull) == 0ull;
        if (__syncthreads_or(need ? 1 : 0)) tab = stage_table<T, kLds>(gtab, a.table_bytes, lds_raw);
    }
    if (i >= a.count || a.done[i]) return;
    const unsigned long long key = a.first_event[i];
    if (key == kNoEvent) return;  // consistent chain and no accept: (0, 0)
    const int64_t ib = a.off[i], ie = a.off[i + 1];
    int64_t g = static_cast<int64_t>(key >> 1);
    if (key & 1ull) {
        a.flags[i] = 1;
        a.lens[i] = static_cast<int32_t>(g * a.seg + a.fixed_acc[g] - ib + 1);
        return;
    }
    // the entry pass 2 assumed for segment g is wrong: continue sequentially from there
    const T *ckpt = static_cast<const T *>(a.ckpt);
    uint32_t cur = ib >= (g - 1) * a.seg ? a.tail_exit[g - 1] : a.fixed_exit[g - 1];
    const int64_t g1 = (ie - 1) / a.seg;
    for (; g <= g1; ++g) {
        const int64_t s0 = g * a.seg;
        int32_t acc_at;
        uint32_t ex;
        if (a.fixed_entry[g] == cur) {
            acc_at = a.fixed_acc[g];
            ex = a.fixed_exit[g];
        } else {
            const int64_t hend = (s0 + a.seg) < ie ? (s0 + a.seg) : ie;
            split_merge_walk<T>(tab, a, s0, hend, cur, ckpt + g * a.nck, a.head_exit[g],
                                acc_at, ex);
        }
        if (acc_at >= 0) {
            a.flags[i] = 1;
            a.lens[i] = static_cast<int32_t>(s0 + acc_at - ib + 1);
            return;
        }
        cur = ex;
    }
}

// ---------------------------------------------------------------- host side -------------

static size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

int32_t device_cu_count() {
    static int32_t cached[64] = {0};
    int dev = 0;
    GFSM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) dev = 0;
    if (!cached[dev]) {
        int cus = 0;
        GFSM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        cached[dev] = cus > 0 ? cus : 1;
    }
    return cached[dev];
}

size_t dfa_lds_bytes(int32_t n) { return static_cast<size_t>(n) * kAlphabet * sizeof(uint16_t); }

static bool lds_fits(const DfaDev &d) { return !d.wide && dfa_lds_bytes(d.n) <= 160 * 1024; }

DfaSplitPlan dfa_split_plan(const DfaDev &d, int64_t total_bytes, int32_t count) {
    DfaSplitPlan p;
    if (lds_fits(d)) {
        p.seg_bytes = 1024;
        p.ckpt_shift = 6;
    } else {
        int64_t c = 1024;
        while (c < 4ll * d.n && c < 65536) c <<= 1;
        p.seg_bytes = c;
        p.ckpt_shift = 6;
        while ((1ll << p.ckpt_shift) * 16 < c) ++p.ckpt_shift;
    }
    p.segments = (total_bytes + p.seg_bytes - 1) / p.seg_bytes;
    const size_t S = static_cast<size_t>(p.segments > 0 ? p.segments : 1);
    const size_t nck = static_cast<size_t>(p.seg_bytes >> p.ckpt_shift);
    const size_t ent = d.wide ? 4 : 2;
    p.scratch_bytes = align256(static_cast<size_t>(count) * 4) +
                      align256(static_cast<size_t>(count) * 8) + 5 * align256(S * 4) +
                      align256(S * nck * ent);
    return p;
}

// Raise a kernel's dynamic-LDS limit to `bytes` (needed above 64 KiB). The limit is what
// the launch uses, not the CU maximum: the kernel's own static LDS counts against the CU.
template <typename K>
static void allow_big_lds(K *kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    GFSM_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(bytes)));
}

template <typename T, bool kLds>
static void split_launch_passes(const DfaDev &d, SplitArgs &a, hipStream_t stream) {
    const T *tab = static_cast<const T *>(d.table);
    const size_t lds = kLds ? static_cast<size_t>(a.table_bytes) : 0;
    const int32_t cus = device_cu_count();
    static size_t attr = 0;  // largest table the attributes were raised for
    if (kLds && lds > attr) {
        allow_big_lds(&dfa_split_walk_kernel<T, kLds>, lds);
        allow_big_lds(&dfa_split_fix_kernel<T, kLds>, lds);
        allow_big_lds(&dfa_split_finish_kernel<T, kLds>, lds);
        attr = lds;
    }
    // spread the segments over every CU
    int64_t per_cu = (a.segments + cus - 1) / cus;
    per_cu = ((per_cu + 63) / 64) * 64;
    int32_t threads, blocks;
    if (kLds) {  // the table fills the LDS: one block per CU, up to 16 waves
        threads = static_cast<int32_t>(per_cu < 64 ? 64 : (per_cu > 1024 ? 1024 : per_cu));
        const int64_t need = (a.segments + threads - 1) / threads;
        blocks = static_cast<int32_t>(need < cus ? (need > 0 ? need : 1) : cus);
    } else {
        threads = per_cu >= 256 ? 256 : 64;
        const int64_t need = (a.segments + threads - 1) / threads;
        const int64_t cap = static_cast<int64_t>(cus) * 32;
        blocks = static_cast<int32_t>(need < cap ? (need > 0 ? need : 1) : cap);
    }
    dfa_split_walk_kernel<T, kLds><<<blocks, threads, lds, stream>>>(tab, a);
    GFSM_HIP(hipGetLastError());
    if (a.segments > 1) {
        dfa_split_fix_kernel<T, kLds><<<blocks, threads, lds, stream>>>(tab, a, 0);
        GFSM_HIP(hipGetLastError());
        dfa_split_fix_kernel<T, kLds><<<blocks, threads, lds, stream>>>(tab, a, 1);
        GFSM_HIP(hipGetLastError());
        const int64_t eblocks64 = (a.segments + 255) / 256;
        const int32_t eblocks = static_cast<int32_t>(eblocks64 < 65536 ? eblocks64 : 65536);
        dfa_split_event_kernel<<<eblocks, 256, 0, stream>>>(a);
        GFSM_HIP(hipGetLastError());
        dfa_split_finish_kernel<T, kLds><<<(a.count + 255) / 256, 256, lds, stream>>>(tab, a);
        GFSM_HIP(hipGetLastError());
    }
}

void launch_dfa_split(const DfaDev &d, const DfaSplitPlan &plan, const uint8_t *data,
                      const int64_t *off, int32_t count, int64_t total_bytes, int32_t *flags,
                      int32_t *lens, void *scratch, hipStream_t stream) {
    if (count <= 0) return;
    if (reinterpret_cast<uintptr_t>(data) & 15)
        throw Error(3, "split needs the input buffer 16-byte aligned");
    SplitArgs a{};
    a.data = data;
    a.off = off;
    a.count = count;
    a.total = total_bytes;
    a.seg = plan.seg_bytes;
    a.rshift = plan.ckpt_shift;
    a.nck = plan.seg_bytes >> plan.ckpt_shift;
    a.segments = plan.segments;
    a.start = d.start_entry;
    a.acc = d.accept_from;
    a.col = d.col_bytes;
    a.table_bytes = d.col_bytes * kAlphabet;
    a.flags = flags;
    a.lens = lens;
    const size_t S = static_cast<size_t>(plan.segments > 0 ? plan.segments : 1);
    char *base = static_cast<char *>(scratch);
    a.done = reinterpret_cast<int32_t *>(base);
    base += align256(static_cast<size_t>(count) * 4);
    a.first_event = reinterpret_cast<unsigned long long *>(base);
    base += align256(static_cast<size_t>(count) * 8);
    a.head_exit = reinterpret_cast<uint32_t *>(base);
    base += align256(S * 4);
    a.tail_exit = reinterpret_cast<uint32_t *>(base);
    base += align256(S * 4);
    a.fixed_entry = reinterpret_cast<uint32_t *>(base);
    base += align256(S * 4);
    a.fixed_exit = reinterpret_cast<uint32_t *>(base);
    base += align256(S * 4);
    a.fixed_acc = reinterpret_cast<int32_t *>(base);
    base += align256(S * 4);
    a.ckpt = base;
    split_init_kernel<<<(count + 255) / 256, 256, 0, stream>>>(flags, lens, a.done,
                                                                a.first_event, count);
    GFSM_HIP(hipGetLastError());
    if (plan.segments <= 0) return;
    if (d.wide)
        split_launch_passes<uint32_t, false>(d, a, stream);
    else if (lds_fits(d))
        split_launch_passes<uint16_t, true>(d, a, stream);
    else
        split_launch_passes<uint16_t, false>(d, a, stream);
}

void launch_fill_result(int32_t *flags, int32_t *lens, int32_t count, int32_t flag,
                        int32_t len, hipStream_t stream) {
    if (count <= 0) return;
    fill_result_kernel<<<(count + 255) / 256, 256, 0, stream>>>(flags, lens, count, flag, len);
    GFSM_HIP(hipGetLastError());
}

void launch_dfa_gather(const DfaDev &d, const uint8_t *data, const int64_t *off, int32_t count,
                       int32_t *flags, int32_t *lens, hipStream_t stream) {
    if (count <= 0) return;
    const int grid = (count + 255) / 256;
    const uint32_t tb = d.col_bytes * kAlphabet;
    if (d.wide)
        dfa_walk_kernel<uint32_t, false><<<grid, 256, 0, stream>>>(
            static_cast<const uint32_t *>(d.table), tb, d.col_bytes, d.start_entry,
            d.accept_from, data, off, count, flags, lens);
    else
        dfa_walk_kernel<uint16_t, false><<<grid, 256, 0, stream>>>(
            static_cast<const uint16_t *>(d.table), tb, d.col_bytes, d.start_entry,
            d.accept_from, data, off, count, flags, lens);
    GFSM_HIP(hipGetLastError());
}

void launch_dfa_lds(const DfaDev &d, const uint8_t *data, const int64_t *off, int32_t count,
                    int32_t *flags, int32_t *lens, hipStream_t stream) {
    if (count <= 0) return;
    if (!lds_fits(d)) throw Error(3, "lds: the table does not fit in LDS");
    const uint32_t tb = d.col_bytes * kAlphabet;
    static size_t attr = 0;
    if (tb > attr) {
        allow_big_lds(&dfa_walk_kernel<uint16_t, true>, tb);
        attr = tb;
    }
    int grid = (count + 255) / 256;
    const int cap = 2 * device_cu_count();
    if (grid > cap) grid = cap;
    dfa_walk_kernel<uint16_t, true><<<grid, 256, tb, stream>>>(
        static_cast<const uint16_t *>(d.table), tb, d.col_bytes, d.start_entry, d.accept_from,
        data, off, count, flags, lens);
    GFSM_HIP(hipGetLastError());
}

}  // namespace gfsm
