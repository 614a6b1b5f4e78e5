// Sparse chain technique for chain-structured NFAs (pattern unions) on gfx950 (wave64).
//
// Same automata as the tiled bit-parallel technique in nfa.hip (every non-start state has
// at most one non-start predecessor and successor, no epsilon edges, homogeneous labels),
// same slot numbering (4096-state tiles, successor of slot s is s + 1), same semantics
// (`src/gpufsm/reference.py:36-65`: latch-first-match, an empty set ends the input with
// (0, 0)). The tiled pass pays for every state on every byte; here the work follows the
// active set, which in a pattern union is a few dozen of tens of thousands of states.
//
// Work items. Persistent waves take (input, chunk) items from an atomic counter. Long
// inputs are cut into chunks when the start state loops on every byte: every non-start
// state is then at most max_chain - 1 bytes from the start, so a chunk started
// max_chain - 1 bytes early from the start state alone has the exact active set from its
// first byte on; accepts inside that overlap belong to the previous chunk and are ignored.
// An input's verdict is the minimum first-accept position over its chunks (atomicMin on
// best[]), which also lets later chunks of an input with a known earlier match stop.
//
// Depth-1 states are never listed. The first positions entered on byte p only matter for
// their accept bit (bflags), for emptiness, and for their successors, which the next byte c
// adds from pair_slot[pair[p][c]]: the second positions whose first class holds p and whose
// own class holds c, with one accept bit for the whole list. Listed states continue to
// s + 1 on c when bit s + 1 of next[c][tile] is set: one 32-bit load per listed state,
// issued together with the load of its accept word. A chain state has one predecessor, so
// no deduplication is needed (a state that is both a first position and a chain successor
// may be listed twice, which changes no verdict).
//
// Three tiers, each run only on the inputs the previous one could not finish:
//   reg  — the listed states live one per lane in a register, at most 64 per wave. Lanes
//          whose state dies free up; the additions of the next byte go to the lanes free at
//          its start, in lane order (mbcnt over the free mask), so there is no compaction
//          and no LDS. The pair entry and byte flags of the next byte are fetched while the
//          current byte is stepped. An input that needs more than 64 lanes is flagged.
//   lds  — the list lives in LDS (up to 1024 states per wave), compacted by ballot + mbcnt
//          each byte. Also the only tier when the pair lists would be too large; it then
//          adds the start state's successors from per-byte lists of first positions.
//   tile — inputs that overflow the LDS list are run by the tiled bit-parallel pass.
// Upstream has no counterpart: its chain-shaped automata go through the bit-packed
// multistream kernels (`native/bitpacked.cu:12-166`), i.e. the tile tier.
#include "internal.hpp"

#include <climits>

namespace gfsm {

namespace {

constexpr int kChainWaves = 4;  // waves per block (all chain kernels)
// The register tiers are held to 8 waves per SIMD (amdgpu_waves_per_eu, at most 64 VGPRs):
// their step waits on a dependent load per byte, and an allocation a few registers past 64
// would cost an eighth of the waves that hide it.

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int32_t lane_rank(uint64_t m) {
    return static_cast<int32_t>(__builtin_amdgcn_mbcnt_hi(
        static_cast<uint32_t>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u)));
}

// Rank of the lane among the set lanes of m in its half of the wave: lo_keep is ~0 for
// lanes 0-31 and 0 for lanes 32-63 (whose low-half bits all count as below them).
__device__ __forceinline__ int32_t lane_rank_half(uint64_t m, uint32_t lo_keep) {
    return static_cast<int32_t>(__builtin_amdgcn_mbcnt_hi(
        static_cast<uint32_t>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m) & lo_keep, 0u)));
}

// Scratch of one launch. Work items are handed out chunk-major: round k holds chunk k of
// every input that has more than k chunks, so the chunks of one input run one round apart
// and a later chunk usually starts after an earlier one has found the input's first match
// (it then skips at once). Inputs are ordered by chunk count, descending (a counting sort),
// which makes round k a prefix of that order: item x of round k is perm[x - rbase[k]].
struct ChainScratch {
    int32_t *best;             // [count] first accept (1-based), INT_MAX = none yet
    int32_t *ovf2;             // [count] inputs the LDS tier hands to the tiles
    int32_t *ovf_items;        // items the register tier hands to the LDS tier
    int32_t *deferred;         // [count] an item of the input overflowed the register tier:
                               // its later items go to the LDS tier without a register run
    int32_t *nch;              // [count] chunks per input
    int32_t *perm;             // [count] inputs by chunk count, descending
    int32_t *hist;             // [K] inputs with exactly c chunks
    int32_t *cursor;           // [K] scatter cursors
    int32_t *rcount;           // [K] R_c = inputs with more than c chunks
    int64_t *rbase;            // [K] first item of round k; total items at [maxk]
    int32_t *maxk;             // [1] most chunks of any input
    unsigned int *counter;     // [4]: reg tier, lds tier, ovf_items fill, inputs in ovf2
    int32_t K;                 // capacity of the per-round arrays (> max chunks + 1)
};

__global__ void chain_zero_kernel(ChainScratch sc) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < sc.K) {
        sc.hist[i] = 0;
        sc.cursor[i] = 0;
    }
    if (i == 0) {
        *sc.maxk = 0;
        sc.counter[0] = 0u;
        sc.counter[1] = 0u;
        sc.counter[2] = 0u;
        sc.counter[3] = 0u;
    }
}

// Chunk counts below kBins are counted per block in LDS first: most inputs have one or two
// chunks, and one global atomic per input on those few counters serializes the kernel.
constexpr int kBins = 256;

__global__ __launch_bounds__(256) void chain_prep_kernel(const int64_t *__restrict__ off,
                                                         int32_t count, int64_t chunk,
                                                         ChainScratch sc) {
    __shared__ int32_t h[kBins];
    __shared__ int32_t bmax;
    for (int k = threadIdx.x; k < kBins; k += blockDim.x) h[k] = 0;
    if (threadIdx.x == 0) bmax = 0;
    __syncthreads();
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        sc.best[i] = INT_MAX;
        sc.ovf2[i] = 0;
        sc.deferred[i] = 0;
        const int64_t len = off[i + 1] - off[i];
        const int32_t c = static_cast<int32_t>((len + chunk - 1) / chunk);
        sc.nch[i] = c;
        if (c > 0) {
            if (c < kBins) atomicAdd(&h[c], 1);
            else atomicAdd(&sc.hist[c], 1);
            atomicMax(&bmax, c);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kBins; k += blockDim.x)
        if (h[k]) atomicAdd(&sc.hist[k], h[k]);
    if (threadIdx.x == 0 && bmax > 0) atomicMax(sc.maxk, bmax);
}

// One block: R_c = sum of hist over (c, maxk] (a suffix sum), then rbase = exclusive prefix
// sums of R over rounds 0..maxk-1 (rbase[maxk] = number of items).
__global__ __launch_bounds__(1024) void chain_rounds_kernel(ChainScratch sc) {
    __shared__ int64_t part[1024];
    const int32_t maxk = *sc.maxk;
    const int32_t t = static_cast<int32_t>(threadIdx.x);
    int64_t carry = 0;
    // suffix sums: position p of the walk is c = maxk - p
    for (int32_t base = 0; base <= maxk; base += 1024) {
        const int32_t c = maxk - (base + t);
        const int64_t v = (c >= 1) ? sc.hist[c] : 0;
        part[t] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int64_t add = t >= d ? part[t - d] : 0;
            __syncthreads();
            part[t] += add;
            __syncthreads();
        }
        // inclusive sum over [c, maxk] minus hist[c] = R_c
        if (c >= 0) sc.rcount[c] = static_cast<int32_t>(carry + part[t] - v);
        const int64_t tile_total = part[1023];
        __syncthreads();
        carry += tile_total;
    }
    __syncthreads();
    carry = 0;
    for (int32_t base = 0; base <= maxk; base += 1024) {
        const int32_t k = base + t;
        const int64_t v = k < maxk ? sc.rcount[k] : 0;
        part[t] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int64_t add = t >= d ? part[t - d] : 0;
            __syncthreads();
            part[t] += add;
            __syncthreads();
        }
        if (k <= maxk) sc.rbase[k] = carry + part[t] - v;
        const int64_t tile_total = part[1023];
        __syncthreads();
        carry += tile_total;
    }
}

// perm[R_c + j] = the j-th input with exactly c chunks (any order within c). Ranks within
// a block come from LDS counters; one global atomic per (block, chunk count) reserves them.
__global__ __launch_bounds__(256) void chain_scatter_kernel(ChainScratch sc, int32_t count) {
    __shared__ int32_t cnt[kBins], base[kBins];
    for (int k = threadIdx.x; k < kBins; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t c = i < count ? sc.nch[i] : 0;
    int32_t local = 0;
    if (c > 0 && c < kBins) local = atomicAdd(&cnt[c], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < kBins; k += blockDim.x)
        if (cnt[k]) base[k] = atomicAdd(&sc.cursor[k], cnt[k]);
    __syncthreads();
    if (c > 0) {
        const int32_t pos = c < kBins ? base[c] + local : atomicAdd(&sc.cursor[c], 1);
        sc.perm[sc.rcount[c] + pos] = i;
    }
}

// Bytes [a, a + 4) of the batch as one little-endian word (a is 4-aligned in absolute
// address terms); bytes outside [0, data_end) read as zero, words at or past stop as zero.
__device__ __forceinline__ uint32_t chain_load_word(const uint8_t *__restrict__ data, int64_t a,
                                                    int64_t stop, int64_t data_end) {
    if (a >= stop) return 0;
    // streamed once: non-temporal, so it does not push the automaton tables out of L2
    if (a >= 0 && a + 4 <= data_end)
        return __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(data + a));
    uint32_t w = 0;
    for (int q = 0; q < 4; ++q)
        if (a + q >= 0 && a + q < data_end) w |= static_cast<uint32_t>(data[a + q]) << (8 * q);
    return w;
}

// Byte r (0..255) of the 256-byte window whose 4-byte words are spread over the lanes.
__device__ __forceinline__ uint32_t window_byte(uint32_t word, int32_t r) {
    return (static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(word), r >> 2)) >>
            ((r & 3) << 3)) & 0xffu;
}

// One work item, decoded.
struct Item {
    int32_t in;
    uint32_t id;  // the item's number (round-major order)
    int64_t ib, cs, ce, ps;
};

// Item number `item` (< items) decoded: its round k is the largest with rbase[k] <= item.
__device__ __forceinline__ void decode_item(uint32_t item, const ChainScratch &sc,
                                            const int64_t *__restrict__ off, int32_t maxk,
                                            int64_t chunk, int64_t overlap, Item &it) {
    int32_t lo = 0, hi = maxk;
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (sc.rbase[mid] <= static_cast<int64_t>(item)) lo = mid; else hi = mid;
    }
    const int32_t in = sc.perm[static_cast<int64_t>(item) - sc.rbase[lo]];
    it.in = in;
    it.id = item;
    it.ib = off[in];
    const int64_t ie = off[in + 1];
    it.cs = it.ib + static_cast<int64_t>(lo) * chunk;
    it.ce = (it.cs + chunk) < ie ? (it.cs + chunk) : ie;
    it.ps = (it.cs - overlap) > it.ib ? (it.cs - overlap) : it.ib;
}

// A wave's block of items taken from a counter: [next, end). Wave-uniform.
struct ItemPool {
    uint32_t next = 0, end = 0;
};

// Items a wave takes per atomic: 1 unless the list is long against the waves in the grid.
// Each hand-out is one atomic on one address, and those serialize (about 11 ns each), so a
// list of many short items (items of settled inputs are skipped at once) costs that much per
// item however little each does. Blocks of up to 16 divide that cost; a list of a few items
// per wave keeps single hand-out, so long items do not queue behind each other.
__device__ __forceinline__ uint32_t pool_block(int64_t items, int64_t waves) {
    const int64_t q = items / (waves * 4);
    return q < 1 ? 1u : (q > 16 ? 16u : static_cast<uint32_t>(q));
}

// Takes the next item for the whole wave; false when none are left. Items are numbered
// 0..n-1, or are list[0..n-1] when a list is given. A wave whose block is used up takes the
// next `per` items with one atomic: every lane takes part (lane 0 adds `per`, the others 0)
// and the value is read from lane 0 under a full mask, so no branch precedes the broadcast.
// (A lane-0-only branch there lets the compiler peel lane 0 off into a loop of its own,
// after which the other lanes never see a new item.) The counter ends above n by at most
// `per` per wave.
__device__ __forceinline__ bool next_item(unsigned int *counter, uint32_t per, ItemPool &pool,
                                          const int32_t *list, const ChainScratch &sc,
                                          const int64_t *__restrict__ off, int32_t maxk, int64_t n,
                                          int64_t chunk, int64_t overlap, int lane, Item &it) {
    if (pool.next == pool.end) {
        const unsigned int mine = atomicAdd(counter, lane == 0 ? per : 0u);
        pool.next = __builtin_amdgcn_readfirstlane(mine);
        pool.end = pool.next + per;
    }
    const uint32_t k = pool.next++;
    if (static_cast<int64_t>(k) >= n) return false;
    const uint32_t item = list ? static_cast<uint32_t>(list[k]) : k;
    decode_item(item, sc, off, maxk, chunk, overlap, it);
    return true;
}

// An earlier first-accept than anything at or after position `from` is already known.
__device__ __forceinline__ bool settled(const int32_t *best, const Item &it, int64_t from) {
    const int32_t bst = __atomic_load_n(&best[it.in], __ATOMIC_RELAXED);
    return bst != INT_MAX && it.ib + bst - 1 < from;
}

// ------------------------------------------------------------------------- reg tier ----

// "Small" automata (at most 16 tiles, 65536 slots): accept bits staged in LDS and 16-bit
// pair lists, which keeps the continuation table and the lists in each XCD's L2 together.
constexpr int kSmallTiles = 16;

// Register tier. Per 256-byte window, each lane holds 4 bytes of input (for the loads) and,
// for the bytes r = 64 q + lane (q = 0..3), everything about byte r that does not depend on
// the listed states: the byte, the pair entry of the depth-1 set before it, whether the
// start state's side accepts on it (a depth-1 accept, or the pair list's accept bit), and
// whether the start state's side is empty after it. The depth-1 set and the start state
// follow from the bytes alone, so this is computed for the whole window up front. A byte
// step then reads three values with readlanes, issues its (at most three) loads together,
// and has one branch, taken only on an event (an accept, an empty set, a list overflow).
template <bool kAlways, bool kSmall>
__global__ __launch_bounds__(64 * kChainWaves) __attribute__((amdgpu_waves_per_eu(8, 8))) void nfa_chain_reg_kernel(
    NfaChainDev d, const uint8_t *__restrict__ data, const int64_t *__restrict__ off,
    int32_t count, ChainScratch sc, int64_t chunk) {
    __shared__ uint32_t s_flags[256];
    __shared__ uint32_t s_acc[kSmall ? (kSmallTiles + 1) * kTileWords * 2 : 1];
    __shared__ int32_t s_defer[kChainWaves][64];  // deferred items, appended 64 at a time
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s_flags[i] = d.bflags[i];
    if (kSmall) {
        const uint32_t *acc32 = reinterpret_cast<const uint32_t *>(d.acc);
        for (int i = threadIdx.x; i < (d.tiles + 1) * kTileWords * 2; i += blockDim.x)
            s_acc[i] = acc32[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int32_t maxk = *sc.maxk;
    const int64_t items = sc.rbase[maxk];
    const int64_t data_end = off[count];
    const int64_t overlap = d.max_chain > 0 ? d.max_chain - 1 : 0;
    int32_t *dbuf = s_defer[threadIdx.x >> 6];
    uint32_t dcount = 0;  // wave-uniform: items staged in dbuf
    // appends the staged items with one atomic (every lane takes part, lane 0 adds)
    auto flush_deferred = [&]() {
        const unsigned int mine = atomicAdd(&sc.counter[2], lane == 0 ? dcount : 0u);
        const unsigned int base = __builtin_amdgcn_readfirstlane(mine);
        if (static_cast<uint32_t>(lane) < dcount) sc.ovf_items[base + lane] = dbuf[lane];
        dcount = 0;
    };
    const char *__restrict__ next_base = reinterpret_cast<const char *>(d.next);
    const uint32_t plane = static_cast<uint32_t>(d.tiles + 1) * (kTileWords * 8);  // bytes per byte plane
    const char *__restrict__ acc_base = reinterpret_cast<const char *>(d.acc);
    const char *__restrict__ slot_base = kSmall ? reinterpret_cast<const char *>(d.pair_slot16)
                                                : reinterpret_cast<const char *>(d.pair_slot);
    Item it;
    ItemPool pool;
    const uint32_t per = pool_block(items, static_cast<int64_t>(gridDim.x) * kChainWaves);
    while (next_item(&sc.counter[0], per, pool, nullptr, sc, off, maxk, items, chunk, overlap,
                     lane, it)) {
        bool stop = settled(sc.best, it, it.cs);  // wave-uniform throughout
        // an item of this input overflowed already: this one goes to the LDS tier unrun,
        // staged and appended with others
        if (!stop && __atomic_load_n(&sc.deferred[it.in], __ATOMIC_RELAXED) != 0) {
            if (dcount == 64) flush_deferred();
            if (lane == 0) dbuf[dcount] = static_cast<int32_t>(it.id);
            ++dcount;
            continue;
        }
        int32_t slot = -1;                        // this lane's listed state, -1 = free
        uint64_t occ = 0;                         // lanes holding a state
        bool s0 = true, ovf = false;              // start state active before the window
        int32_t d1 = 256;                         // byte of the depth-1 set; 256 = none
        int32_t found = 0;
        const int64_t abase =
            it.ps - static_cast<int64_t>((reinterpret_cast<uintptr_t>(data) + it.ps) & 3u);
        for (int64_t w0 = abase; w0 < it.ce && !stop; w0 += 256) {
            const uint32_t word = chain_load_word(data, w0 + 4 * lane, it.ce, data_end);
            const int32_t lo = it.ps > w0 ? static_cast<int32_t>(it.ps - w0) : 0;
            const int32_t hi = (it.ce - w0) < 256 ? static_cast<int32_t>(it.ce - w0) : 256;
            const int32_t hit_from = it.cs > w0 ? static_cast<int32_t>(it.cs - w0) : 0;
            uint32_t bq[4], fq[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t wq = static_cast<uint32_t>(
                    __shfl(static_cast<int>(word), 16 * q + (lane >> 2)));
                bq[q] = (wq >> (8 * (lane & 3))) & 0xffu;
                fq[q] = s_flags[bq[q]];
            }
            // kpos: first byte of the window (>= lo) on which the start state has no
            // self-loop; the start state is active before byte r while s0 && kpos >= r
            int32_t kpos = 1 << 20;
            if (!kAlways) {
#pragma unroll
                for (int q = 3; q >= 0; --q) {
                    const int32_t r = 64 * q + lane;
                    const uint64_t m = __ballot(r >= lo && r < hi && !(fq[q] & 4u));
                    if (m) kpos = 64 * q + static_cast<int32_t>(__builtin_ctzll(m));
                }
            }
            uint32_t info[4];
            int32_t pe_off[4], pe_cnt[4];
            int32_t d1_after = 256;  // lane 63, q = 3: state of the start side after byte 255
            bool s0_after = true;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int32_t r = 64 * q + lane;
                uint32_t bp = static_cast<uint32_t>(__shfl_up(static_cast<int>(bq[q]), 1));
                uint32_t fp = static_cast<uint32_t>(__shfl_up(static_cast<int>(fq[q]), 1));
                if (q > 0 && lane == 0) {
                    bp = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(bq[q - 1]), 63));
                    fp = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(fq[q - 1]), 63));
                }
                const bool s0p = kAlways || (s0 && kpos >= r - 1);  // before byte r - 1
                const bool s0r = kAlways || (s0 && kpos >= r);      // before byte r
                const bool s0n = kAlways || (s0 && kpos > r);       // after byte r
                // r == lo takes the running d1; every later r derives it from byte r - 1
                const int32_t d1q =
                    r == lo ? d1 : ((s0p && (fp & 1u)) ? static_cast<int32_t>(bp) : 256);
                const int32_t d1n = (s0r && (fq[q] & 1u)) ? static_cast<int32_t>(bq[q]) : 256;
                const int32_t idx =
                    (r >= lo && r < hi) ? 2 * (d1q * 256 + static_cast<int32_t>(bq[q])) : 0;
                const int32_t pw = d.pair[idx + 1];
                pe_off[q] = d.pair[idx];
                pe_cnt[q] = pw & INT_MAX;
                const uint32_t start_acc = (pw < 0 || (s0r && (fq[q] & 2u))) ? 1u : 0u;
                const uint32_t start_dead = (d1n == 256 && !s0n) ? 1u : 0u;
                info[q] = bq[q] | (start_acc << 8) | (start_dead << 9);
                if (q == 3) {
                    d1_after = d1n;
                    s0_after = s0n;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int32_t ilo = lo - 64 * q > 0 ? lo - 64 * q : 0;
                const int32_t ihi = hi - 64 * q < 64 ? hi - 64 * q : 64;
                for (int32_t i = ilo; i < ihi && !stop; ++i) {
                    const uint32_t inf = static_cast<uint32_t>(
                        __builtin_amdgcn_readlane(static_cast<int>(info[q]), i));
                    const uint32_t b = inf & 0xffu;
                    const uint32_t add_off = static_cast<uint32_t>(__builtin_amdgcn_readlane(pe_off[q], i));
                    const int32_t add_cnt = __builtin_amdgcn_readlane(pe_cnt[q], i);
                    // additions go to the lanes free at the start of the byte, in lane order
                    const bool have = slot >= 0;
                    const int32_t frank = lane_rank(~occ);
                    int32_t cand = -1;
                    if (!have && frank < add_cnt) {
                        const uint32_t k = add_off + static_cast<uint32_t>(frank);
                        cand = kSmall ? static_cast<int32_t>(
                                            *reinterpret_cast<const uint16_t *>(slot_base + (k << 1)))
                                      : *reinterpret_cast<const int32_t *>(slot_base + (k << 2));
                    }
                    // continuation of held states; free lanes look up slot 0 and ignore it
                    const uint32_t t = have ? static_cast<uint32_t>(slot + 1) : 0u;
                    // plane b of the continuation table, 32-bit word t >> 5
                    const uint32_t noff = b * plane + ((t >> 5) << 2);
                    const uint32_t nw = *reinterpret_cast<const uint32_t *>(next_base + noff);
                    const uint32_t aw = kSmall ? s_acc[t >> 5]
                                               : *reinterpret_cast<const uint32_t *>(acc_base + ((t >> 5) << 2));
                    const uint32_t bit = t & 31u;
                    const bool alive = have & static_cast<bool>((nw >> bit) & 1u);
                    const bool hit = alive & static_cast<bool>((aw >> bit) & 1u);
                    const uint64_t occ_prev = occ;
                    slot = alive ? static_cast<int32_t>(t) : cand;
                    occ = __ballot(slot >= 0);
                    const uint64_t hitm = __ballot(hit);
                    const bool ovf_now = add_cnt > 64 - __popcll(occ_prev);
                    const bool acc_now = hitm != 0 || (inf & 0x100u);
                    const bool dead = occ == 0 && (inf & 0x200u);
                    if (ovf_now || acc_now || dead) {
                        const int32_t r = 64 * q + i;
                        if (ovf_now) {
                            ovf = true;
                            stop = true;
                        } else if (acc_now && r >= hit_from) {
                            found = static_cast<int32_t>(w0 + r - it.ib + 1);
                            stop = true;
                        } else if (dead) {
                            stop = true;
                        }
                    }
                }
            }
            // the start side after the window's last byte (only full windows continue)
            d1 = __builtin_amdgcn_readlane(d1_after, 63);
            s0 = kAlways || (__builtin_amdgcn_readlane(s0_after ? 1 : 0, 63) != 0);
            // another chunk of this input may have found an earlier match meanwhile
            if (!stop && settled(sc.best, it, w0 + 256)) stop = true;
        }
        if (lane == 0) {
            if (found > 0) atomicMin(&sc.best[it.in], found);
            if (ovf) {
                sc.ovf_items[atomicAdd(&sc.counter[2], 1u)] = static_cast<int32_t>(it.id);
                __atomic_store_n(&sc.deferred[it.in], 1, __ATOMIC_RELAXED);
            }
        }
    }
    if (dcount != 0) flush_deferred();  // wave-uniform; the LDS tier launches after this kernel
}

// ---------------------------------------------------------------- paired reg tier ----

// Register tier with two work items per wave: lanes 0-31 run one (input, chunk) item and
// lanes 32-63 another, each half holding up to 32 listed states, one per lane. The step is
// the one-item tier's, but what was uniform over the wave is now uniform over a half: the
// per-byte values (pair list offset, byte, clipped pair count, start-side accept, in range)
// are written to LDS once per window and read back with one broadcast load per byte, ranks
// among a half's free lanes come from mbcnt over that half's bits, and the events of both
// halves (accept, overflow) are caught by one ballot. The per-item scalar work of the
// one-item tier becomes vector work that serves two items per instruction.
//
// A half whose item stops inside a window (an accept, an overflow) clears the in-range bit
// of its remaining bytes and idles to the window's end. An empty active set (possible only
// without the start state's self-loop on every byte) is detected at window ends: the set
// stays empty once the start state and its successors are gone.
constexpr int kHalf = 32;                   // lanes per item
constexpr int kHalfWin = 4 * kHalf;         // bytes per window and item
constexpr uint32_t kInRange = 0x80000000u;  // packed byte info: bit 31
constexpr uint32_t kStartAcc = 0x80u;       // bit 7; bits 0-6: pair count, clipped to 127
// bits 9-28: the byte's plane of the continuation table in 32-bit words when the automaton
// is small (byte * (tiles + 1) * 128 < 2^20), else the byte itself
constexpr int kPlaneShift = 9;

template <bool kAlways, bool kSmall>
__global__ __launch_bounds__(64 * kChainWaves) __attribute__((amdgpu_waves_per_eu(8, 8))) void nfa_chain_pair_kernel(
    NfaChainDev d, const uint8_t *__restrict__ data, const int64_t *__restrict__ off,
    int32_t count, ChainScratch sc, int64_t chunk) {
    __shared__ uint32_t s_flags[256];
    __shared__ uint32_t s_acc[kSmall ? (kSmallTiles + 1) * kTileWords * 2 : 1];
    // {pair list offset, packed byte info} per byte; one spare entry per half, read by the
    // prefetch after the window's last byte
    __shared__ uint2 s_info[kChainWaves][2][kHalfWin + 1];
    // deferred items of each wave, appended to the overflow list 64 at a time: one append
    // atomic per item serializes like the hand-out does
    __shared__ int32_t s_defer[kChainWaves][64];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s_flags[i] = d.bflags[i];
    if (kSmall) {
        const uint32_t *acc32 = reinterpret_cast<const uint32_t *>(d.acc);
        for (int i = threadIdx.x; i < (d.tiles + 1) * kTileWords * 2; i += blockDim.x)
            s_acc[i] = acc32[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    const int hl = lane & 31;      // lane within the half
    const int hbase = lane & 32;   // first lane of the half
    const uint32_t lo_keep = half ? 0u : ~0u;
    uint2 *info = s_info[threadIdx.x >> 6][half];
    const int32_t maxk = *sc.maxk;
    const int64_t items = sc.rbase[maxk];
    const int64_t data_end = off[count];
    const int64_t overlap = d.max_chain > 0 ? d.max_chain - 1 : 0;
    const uint32_t plane_words = static_cast<uint32_t>(d.tiles + 1) * (kTileWords * 2);
    const char *__restrict__ next_base = reinterpret_cast<const char *>(d.next);
    const char *__restrict__ acc_base = reinterpret_cast<const char *>(d.acc);
    const char *__restrict__ slot_base = kSmall ? reinterpret_cast<const char *>(d.pair_slot16)
                                                : reinterpret_cast<const char *>(d.pair_slot);
    // item state, uniform within a half
    Item it{0, 0u, 0, 0, 0, 0};
    bool running = false;    // the half holds an item
    bool exhausted = false;  // wave-uniform: every item has been handed out
    int64_t w0 = 0;          // first byte of the item's current window
    int32_t d1 = 256;        // byte of the depth-1 set before the window; 256 = none
    bool s0 = true;
    // Stop state as integers, not flags: flags changed in divergent code inside the byte
    // loop become lane masks that the compiler merges with exec on every iteration.
    int32_t stopped = 0;     // nonzero once the item is finished
    int32_t found = 0;       // > 0: first accept (1-based); -1: overflow to the LDS tier
    int32_t slot = -1;       // this lane's listed state, -1 = free
    // the first item of every half is its own index; later ones come from the counter,
    // offset by the number of halves (a single counter handed out to every half at once
    // serializes the start of the launch)
    const uint32_t halves = gridDim.x * (kChainWaves * 2);
    const uint32_t first_item = (blockIdx.x * kChainWaves + (threadIdx.x >> 6)) * 2 + half;
    bool first_round = true;  // wave-uniform
    // later items come from the wave's block of counter values (ItemPool), one half served
    // per pass of the loop below
    ItemPool pool;
    const uint32_t per = pool_block(items, static_cast<int64_t>(gridDim.x) * kChainWaves);
    int32_t *dbuf = s_defer[threadIdx.x >> 6];
    uint32_t dcount = 0;  // wave-uniform: items staged in dbuf
    // appends the staged items with one atomic (every lane takes part, lane 0 adds)
    auto flush_deferred = [&]() {
        const unsigned int mine = atomicAdd(&sc.counter[2], lane == 0 ? dcount : 0u);
        const unsigned int base = __builtin_amdgcn_readfirstlane(mine);
        if (static_cast<uint32_t>(lane) < dcount) sc.ovf_items[base + lane] = dbuf[lane];
        dcount = 0;
    };
    for (;;) {
        // halves without an item take the next one (all lanes take part in the atomic and
        // the broadcast; items already settled are skipped here)
        while (!exhausted && __ballot(!running) != 0) {
            if (dcount > 62) flush_deferred();  // room for both halves' items
            uint32_t item = first_item;
            bool take = !running;  // this half takes an item in this pass
            if (!first_round) {
                // the lower half that needs an item takes the pool's next one
                const uint64_t need = __ballot(!running);
                const int taker = (need & 0xffffffffull) ? 0 : 1;  // wave-uniform
                take = !running && half == taker;
                if (pool.next == pool.end) {
                    const unsigned int mine = atomicAdd(&sc.counter[0], lane == 0 ? per : 0u);
                    pool.next = __builtin_amdgcn_readfirstlane(mine);
                    pool.end = pool.next + per;
                }
                item = pool.next + halves;
                ++pool.next;
            }
            first_round = false;
            bool none = false;
            bool defer = false;  // the item taken goes to the LDS tier unrun
            if (take) {
                if (static_cast<int64_t>(item) < items) {
                    decode_item(item, sc, off, maxk, chunk, overlap, it);
                    const bool skip = settled(sc.best, it, it.cs);
                    // an item of this input overflowed already (a dense input overflows in
                    // every chunk): this one goes to the LDS tier unrun
                    defer = !skip && __atomic_load_n(&sc.deferred[it.in], __ATOMIC_RELAXED) != 0;
                    if (!skip && !defer) {
                        running = true;
                        w0 = it.ps - static_cast<int64_t>(
                                         (reinterpret_cast<uintptr_t>(data) + it.ps) & 3u);
                        d1 = 256;
                        s0 = true;
                        stopped = 0;
                        found = 0;
                        slot = -1;
                    }
                } else {
                    none = true;
                }
            }
            // stage deferred items: lane 0 of each deferring half, the lower half first
            const uint64_t dm = __ballot(defer && hl == 0);
            if (defer && hl == 0)
                dbuf[dcount + ((half && (dm & 1ull)) ? 1u : 0u)] = static_cast<int32_t>(it.id);
            dcount += static_cast<uint32_t>(__popcll(dm));
            exhausted = __ballot(none) != 0;
        }
        if (__ballot(running) == 0) break;  // nothing left to hand out and nothing running

        // ---- the window: per-byte values of both halves into LDS
        const uint32_t word = running ? chain_load_word(data, w0 + 4 * hl, it.ce, data_end) : 0u;
        const int32_t lo = running ? (it.ps > w0 ? static_cast<int32_t>(it.ps - w0) : 0) : kHalfWin;
        const int32_t hi =
            running ? ((it.ce - w0) < kHalfWin ? static_cast<int32_t>(it.ce - w0) : kHalfWin) : 0;
        const int32_t hit_from = it.cs > w0 ? static_cast<int32_t>(it.cs - w0) : 0;
        uint32_t bq[4], fq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t wq = static_cast<uint32_t>(
                __shfl(static_cast<int>(word), hbase + 8 * q + (hl >> 2)));
            bq[q] = (wq >> (8 * (hl & 3))) & 0xffu;
            fq[q] = s_flags[bq[q]];
        }
        // kpos: first byte of the window (>= lo) on which the start state has no self-loop
        int32_t kpos = 1 << 20;
        if (!kAlways) {
#pragma unroll
            for (int q = 3; q >= 0; --q) {
                const int32_t r = kHalf * q + hl;
                const uint64_t m = __ballot(r >= lo && r < hi && !(fq[q] & 4u));
                const uint32_t mh = static_cast<uint32_t>(m >> hbase);
                if (mh) kpos = kHalf * q + static_cast<int32_t>(__builtin_ctz(mh));
            }
        }
        int32_t d1_after = 256;  // hl 31, q 3: the start side after the window's last byte
        bool s0_after = true;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t r = kHalf * q + hl;
            uint32_t bp = static_cast<uint32_t>(__shfl_up(static_cast<int>(bq[q]), 1, kHalf));
            uint32_t fp = static_cast<uint32_t>(__shfl_up(static_cast<int>(fq[q]), 1, kHalf));
            if (q > 0) {
                const uint32_t cb = static_cast<uint32_t>(__shfl(static_cast<int>(bq[q - 1]), hbase + 31));
                const uint32_t cf = static_cast<uint32_t>(__shfl(static_cast<int>(fq[q - 1]), hbase + 31));
                if (hl == 0) {
                    bp = cb;
                    fp = cf;
                }
            }
            const bool s0p = kAlways || (s0 && kpos >= r - 1);  // before byte r - 1
            const bool s0r = kAlways || (s0 && kpos >= r);      // before byte r
            const bool s0n = kAlways || (s0 && kpos > r);       // after byte r
            const int32_t d1q =
                r == lo ? d1 : ((s0p && (fp & 1u)) ? static_cast<int32_t>(bp) : 256);
            const int32_t d1n = (s0r && (fq[q] & 1u)) ? static_cast<int32_t>(bq[q]) : 256;
            const bool inr = r >= lo && r < hi;
            const int32_t idx = inr ? 2 * (d1q * 256 + static_cast<int32_t>(bq[q])) : 0;
            const int32_t pw = d.pair[idx + 1];
            const int32_t cnt = pw & INT_MAX;
            const uint32_t sacc = (inr && (pw < 0 || (s0r && (fq[q] & 2u)))) ? kStartAcc : 0u;
            const uint32_t where = kSmall ? bq[q] * plane_words : bq[q];
            const uint32_t pk = (where << kPlaneShift) | sacc |
                                (inr ? (kInRange | static_cast<uint32_t>(cnt < 127 ? cnt : 127)) : 0u);
            // an empty list reads entries 0-31 (one line that stays cached), not the scattered
            // offset of its (depth-1 byte, byte) pair
            const uint32_t po = (inr && cnt > 0) ? static_cast<uint32_t>(d.pair[idx]) : 0u;
            info[r] = make_uint2(po, pk);
            if (q == 3) {
                d1_after = d1n;
                s0_after = s0n;
            }
        }
        wave_sync();

        // ---- the bytes: both halves step together
        const int32_t ilo = min(__builtin_amdgcn_readlane(lo, 0), __builtin_amdgcn_readlane(lo, 32));
        const int32_t ihi = max(__builtin_amdgcn_readlane(hi, 0), __builtin_amdgcn_readlane(hi, 32));
        // word offset of 32-bit word wd of a byte's plane of the continuation table
        auto next_off = [&](uint32_t pk, uint32_t wd) -> uint32_t {
            return kSmall ? (((pk >> kPlaneShift) & 0xfffffu) + wd) << 2
                          : ((pk >> kPlaneShift) & 0xffu) * (plane_words << 2) + (wd << 2);
        };
        auto acc_word = [&](uint32_t wd) -> uint32_t {
            return kSmall ? s_acc[wd] : *reinterpret_cast<const uint32_t *>(acc_base + (wd << 2));
        };
        auto list_entry = [&](uint32_t k) -> int32_t {
            return kSmall ? static_cast<int32_t>(*reinterpret_cast<const uint16_t *>(slot_base + (k << 1)))
                          : *reinterpret_cast<const int32_t *>(slot_base + (k << 2));
        };

        // One byte of both halves; returns true when neither half has an item left to step.
        // The step issues about 30 VALU and 10 SALU instructions per wave and is bound by
        // instruction issue at full occupancy rather than by the latency of its loads:
        // stepping two bytes per memory round trip (the second byte's continuation loaded
        // speculatively, its additions found by cross-lane reads of read-ahead pair lists)
        // was measured 20% slower.
        auto one_byte = [&](int32_t i) -> bool {
            const uint2 e = info[i];
            const uint32_t po = e.x, pk = e.y;
            const bool inr = static_cast<int32_t>(pk) < 0;
            const bool was_free = slot < 0;
            const int32_t frank = lane_rank_half(__ballot(was_free), lo_keep);
            // additions go to the half's free lanes in lane order; every lane reads an entry
            // (at most 31 past the list, inside the padding) and only taking lanes use it
            const int32_t cand = list_entry(po + static_cast<uint32_t>(frank));
            // continuation of held states; free lanes look up slot 0, whose bit is always 0
            const uint32_t t = static_cast<uint32_t>(slot + 1);
            const uint32_t wd = t >> 5;
            const uint32_t nw = *reinterpret_cast<const uint32_t *>(next_base + next_off(pk, wd));
            const uint32_t aw = acc_word(wd);
            // all loads in flight together: without this the compiler moves the pair list
            // load under the "state died" branch, after the continuation load has returned
            asm volatile("" ::"v"(nw), "v"(cand), "v"(aw));
            const uint32_t bit = 1u << (t & 31u);
            const uint32_t alive = nw & bit;
            const bool hit = inr && (alive & aw) != 0;
            const int32_t cnt = static_cast<int32_t>(pk & 127u);
            const bool take = was_free && frank < cnt;
            // lane 31 of a half: its rank plus its own free bit = the half's free lanes
            const int32_t nfree = frank + (was_free ? 1 : 0);
            const bool short_now = hl == 31 && cnt > nfree;
            const int32_t nslot = alive ? static_cast<int32_t>(t) : (take ? cand : -1);
            slot = inr ? nslot : slot;
            if (__ballot(hit || (pk & kStartAcc) || short_now) == 0) return false;
            // more additions than lanes free before the byte (rare): the rest go to the
            // lanes whose state died on this byte, in lane order; only when those are too
            // few as well does the item overflow to the LDS tier
            const uint64_t sm = __ballot(short_now);
            const int32_t F = __shfl(nfree, hbase + 31);
            const bool died = inr && !was_free && !alive;
            const int32_t drank = lane_rank_half(__ballot(died), lo_keep);
            const int32_t ndied = __shfl(drank + (died ? 1 : 0), hbase + 31);
            const bool short_half = static_cast<uint32_t>(sm >> hbase) != 0;
            if (short_half && died && F + drank < cnt)
                slot = list_entry(po + static_cast<uint32_t>(F + drank));
            const bool ovf_half = short_half && cnt > F + ndied;
            const uint64_t em = __ballot(hit || (pk & kStartAcc) || ovf_half);
            if (static_cast<uint32_t>(em >> hbase) != 0) {
                if (ovf_half) {
                    found = -1;
                    stopped = i + 1;
                } else if (i >= hit_from) {
                    found = static_cast<int32_t>(w0 + i - it.ib + 1);
                    stopped = i + 1;
                }
                if (stopped) {
                    slot = -1;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int32_t r = kHalf * q + hl;
                        if (r > i) info[r].y = 0u;
                    }
                }
            }
            wave_sync();
            return __ballot(running && stopped == 0) == 0;
        };

        for (int32_t i = ilo; i < ihi; ++i)
            if (one_byte(i)) break;

        // ---- window end: carry the start side, finish items
        d1 = __shfl(d1_after, hbase + 31);
        s0 = kAlways || __shfl(s0_after ? 1 : 0, hbase + 31) != 0;
        w0 += kHalfWin;
        const bool held = static_cast<uint32_t>(__ballot(slot >= 0) >> hbase) != 0;
        if (running && stopped == 0) {
            if (w0 >= it.ce) stopped = 1;
            else if (!kAlways && !held && d1 == 256 && !s0) stopped = 1;  // empty for good
            else if (settled(sc.best, it, w0)) stopped = 1;
        }
        if (running && stopped != 0) {
            if (hl == 0) {
                if (found > 0) atomicMin(&sc.best[it.in], found);
                if (found < 0) {
                    sc.ovf_items[atomicAdd(&sc.counter[2], 1u)] = static_cast<int32_t>(it.id);
                    __atomic_store_n(&sc.deferred[it.in], 1, __ATOMIC_RELAXED);
                }
            }
            running = false;
            slot = -1;
        }
    }
    if (dcount != 0) flush_deferred();  // wave-uniform; the LDS tier launches after this kernel
}

// ------------------------------------------------------------------------ lane tier ----

// Lane tier: pattern unions whose start state loops on every byte, with byte-triple lists.
//
// With the start state always active, the depth-1 set after byte j is a function of b_j,
// the depth-2 set of (b_j-1, b_j), and every state past that entered at depth 3 from the list
// tri(b_j-2, b_j-1, b_j) (see TripleHost). Pattern unions hold few such states: on the
// 2000-pattern ragged mix about 0.5 enter per byte, a state survives a byte with
// probability about 1/16, and a lane rarely holds more than 3. So each lane runs one
// (input, chunk) item alone, holding up to kTriHeld states in LDS. Per byte it loads one
// continuation word per held state, reads the byte's list (whose header it loaded together
// with those of the other bytes of the 16-byte group), and checks the depth-1, depth-2 and
// depth-3 accepts against the LDS flags, the pair accept bits and the header. No lane
// waits for another, so a wave steps 64 items per byte where the paired register tier steps
// 2, and a byte costs about two scattered loads instead of one per listed state.
// An item whose held set would exceed kTriHeld goes to the LDS tier unrun (like the register
// tiers' overflows), and so do the later items of its input.
constexpr int kTriHeld = 8;
constexpr int kTriWaves = 4;

// Byte k of a 16-byte group.
__device__ __forceinline__ uint32_t group_byte(const uint4 &v, int k) {
    const uint32_t w = k < 4 ? v.x : (k < 8 ? v.y : (k < 12 ? v.z : v.w));
    return (w >> (8 * (k & 3))) & 0xffu;
}

// The 16 bytes at [ga, ga + 16) of the batch, ga 16-aligned in absolute address terms; bytes
// outside [0, data_end) read as zero (byte loads at the batch's two ends).
__device__ __forceinline__ uint4 tri_load_group(const uint8_t *__restrict__ data, int64_t ga,
                                                int64_t data_end) {
    if (ga >= 0 && ga + 16 <= data_end) {
        // streamed once: non-temporal, so it does not push the lists out of L2
        const uint32_t *p = reinterpret_cast<const uint32_t *>(data + ga);
        return make_uint4(__builtin_nontemporal_load(p), __builtin_nontemporal_load(p + 1),
                          __builtin_nontemporal_load(p + 2), __builtin_nontemporal_load(p + 3));
    }
    uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < 16; ++i)
        if (ga + i >= 0 && ga + i < data_end)
            w[i >> 2] |= static_cast<uint32_t>(data[ga + i]) << (8 * (i & 3));
    return make_uint4(w[0], w[1], w[2], w[3]);
}

__global__ __launch_bounds__(64 * kTriWaves) void nfa_chain_tri_kernel(
    NfaChainDev d, NfaTriDev tri, const uint8_t *__restrict__ data,
    const int64_t *__restrict__ off, int32_t count, ChainScratch sc, int64_t chunk) {
    __shared__ uint32_t s_flags[256];
    __shared__ uint32_t s_acc[(kSmallTiles + 1) * kTileWords * 2];
    __shared__ uint32_t s_pacc[2048];
    __shared__ uint4 s_held[64 * kTriWaves];  // kTriHeld 16-bit slots per lane
    __shared__ uint32_t s_hdr[16][64 * kTriWaves];  // the group's list headers, by byte
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s_flags[i] = d.bflags[i];
    const int acc_words = (d.tiles + 1) * kTileWords * 2;
    const uint32_t *acc32 = reinterpret_cast<const uint32_t *>(d.acc);
    for (int i = threadIdx.x; i < acc_words; i += blockDim.x) s_acc[i] = acc32[i];
    for (int i = threadIdx.x; i < 2048; i += blockDim.x) s_pacc[i] = tri.pair_acc[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    uint16_t *held = reinterpret_cast<uint16_t *>(&s_held[threadIdx.x]);
    const int32_t maxk = *sc.maxk;
    const int64_t items = sc.rbase[maxk];
    const int64_t data_end = off[count];
    const int64_t overlap = d.max_chain > 0 ? d.max_chain - 1 : 0;
    const uint32_t plane_words = static_cast<uint32_t>(d.tiles + 1) * (kTileWords * 2);
    const char *__restrict__ next_base = reinterpret_cast<const char *>(d.next);
    const uint64_t dbase = static_cast<uint64_t>(reinterpret_cast<uintptr_t>(data));

    // the lane's item
    Item it{0, 0u, 0, 0, 0, 0};
    bool running = false;
    bool fresh = false;    // the item's first group is next
    int64_t q = 0;         // next byte to step
    int32_t n = 0;         // held states
    uint32_t x = 0, y = 0; // the bytes at q - 2 and q - 1
    int32_t hist = 0;      // bytes of the item stepped so far, capped at 2
    int32_t found = 0;     // > 0: first accept (1-based); -1: overflow to the LDS tier
    int32_t groups = 0;    // groups stepped in this item
    uint4 v = make_uint4(0u, 0u, 0u, 0u);  // bytes of the group at q (16-aligned)

    // Hand-out. Every lane's first item is its own index; later ones come from the counter,
    // offset by the number of lanes, a wave taking kPer numbers per atomic (every lane takes
    // part in the atomic and the broadcast, see next_item).
    constexpr uint32_t kPer = 64;
    const uint32_t lanes_total = gridDim.x * blockDim.x;
    bool first_round = true;  // wave-uniform
    bool exhausted = false;   // wave-uniform
    ItemPool pool;

    for (;;) {
        while (!exhausted) {
            const uint64_t need = __ballot(!running);
            if (need == 0) break;
            uint32_t item;
            bool take;
            if (first_round) {
                item = blockIdx.x * blockDim.x + threadIdx.x;
                take = true;
                first_round = false;
            } else {
                if (pool.next == pool.end) {
                    const unsigned int mine = atomicAdd(&sc.counter[0], lane == 0 ? kPer : 0u);
                    pool.next = __builtin_amdgcn_readfirstlane(mine);
                    pool.end = pool.next + kPer;
                }
                const uint32_t avail = pool.end - pool.next;
                const uint32_t r = static_cast<uint32_t>(lane_rank(need));
                take = !running && r < avail;
                item = pool.next + lanes_total + r;
                const uint32_t k = static_cast<uint32_t>(__popcll(need));
                pool.next += k < avail ? k : avail;
            }
            bool none = false;
            if (take) {
                if (static_cast<int64_t>(item) < items) {
                    decode_item(item, sc, off, maxk, chunk, overlap, it);
                    if (!settled(sc.best, it, it.cs)) {
                        if (__atomic_load_n(&sc.deferred[it.in], __ATOMIC_RELAXED) != 0) {
                            // an item of this input overflowed: the LDS tier runs this one
                            sc.ovf_items[atomicAdd(&sc.counter[2], 1u)] = static_cast<int32_t>(it.id);
                        } else {
                            running = true;
                            fresh = true;
                            q = it.ps;
                            n = 0;
                            hist = 0;
                            found = 0;
                            groups = 0;
                        }
                    }
                } else {
                    none = true;
                }
            }
            if (__ballot(none) != 0) exhausted = true;
        }
        if (__ballot(running) == 0) break;  // nothing left to hand out and nothing running

        // ---- one 16-byte group of every running lane
        const int64_t ga = q - static_cast<int64_t>((dbase + static_cast<uint64_t>(q)) & 15u);
        const int64_t lim = running ? it.ce : ga;  // bytes below lim are stepped
        const int32_t s = static_cast<int32_t>(q - ga);  // the group's first byte to step
        const int32_t e = static_cast<int32_t>(lim - ga < 16 ? lim - ga : 16);  // ... and end
        if (running && fresh) v = tri_load_group(data, ga, data_end);
        fresh = false;
        // the headers of the group's depth-3 lists, kept in LDS by byte: the non-empty bits
        // first (a table that stays in L2), then the headers of the non-empty lists only
        {
            uint32_t idx[16], hdr[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const uint32_t b = group_byte(v, k);
                const uint32_t y1 = k == 0 ? y : group_byte(v, k - 1);
                const uint32_t x1 = k == 0 ? x : (k == 1 ? y : group_byte(v, k - 2));
                const bool ok = running && k >= s && k < e && hist + (k - s) >= 2;
                idx[k] = ok ? (x1 << 16) | (y1 << 8) | b : ~0u;
                hdr[k] = ok ? tri.nonempty[idx[k] >> 5] : 0u;
            }
#pragma unroll
            for (int k = 0; k < 16; ++k)
                hdr[k] = (hdr[k] >> (idx[k] & 31u)) & 1u ? tri.hdr[idx[k]] : 0u;
#pragma unroll
            for (int k = 0; k < 16; ++k) s_hdr[k][threadIdx.x] = hdr[k];
        }
        // the next group's bytes, read while this one is stepped
        const bool more = running && e == 16 && ga + 16 < lim;
        const uint4 vn = more ? tri_load_group(data, ga + 16, data_end) : make_uint4(0u, 0u, 0u, 0u);

        bool stop = false;
        for (int k = 0; k < 16; ++k) {
            const bool act = running && !stop && k >= s && k < e;
            if (__ballot(act) == 0) continue;
            const uint32_t wk = k < 4 ? v.x : (k < 8 ? v.y : (k < 12 ? v.z : v.w));
            const uint32_t b = (wk >> (8 * (k & 3))) & 0xffu;
            const uint32_t wp = k <= 4 ? v.x : (k <= 8 ? v.y : (k <= 12 ? v.z : v.w));
            const uint32_t yb = k == 0 ? y : ((wp >> (8 * ((k - 1) & 3))) & 0xffu);
            const int32_t hk = hist + (k - s);  // item bytes stepped before this one
            // held states: their continuation words on b, loaded together
            uint4 hv = make_uint4(0u, 0u, 0u, 0u);
            if (__ballot(act && n > 0) != 0) hv = s_held[threadIdx.x];
            uint32_t t[kTriHeld], nw[kTriHeld];
#pragma unroll
            for (int i = 0; i < kTriHeld; ++i) {
                const uint32_t hw = i < 2 ? hv.x : (i < 4 ? hv.y : (i < 6 ? hv.z : hv.w));
                t[i] = ((hw >> (16 * (i & 1))) & 0xffffu) + 1u;
                nw[i] = 0u;
                if (__ballot(act && n > i) != 0 && act && n > i)
                    nw[i] = *reinterpret_cast<const uint32_t *>(
                        next_base + ((b * plane_words + (t[i] >> 5)) << 2));
            }
            // the byte's depth-3 list: inline when it has one entry, else 4 entries a load
            const uint32_t h = s_hdr[k][threadIdx.x];
            const uint32_t cnt = act ? (h >> 24) & 127u : 0u;
            uint2 e4 = make_uint2(0u, 0u);
            if (__ballot(cnt >= 2) != 0 && cnt >= 2)
                e4 = *reinterpret_cast<const uint2 *>(tri.slot + (h & 0xffffffu));
            // survivors, compacted to the front of the lane's list (read before written)
            int32_t p = 0;
            bool hit = false;
#pragma unroll
            for (int i = 0; i < kTriHeld; ++i) {
                if (__ballot(act && n > i) != 0) {
                    const bool alive = act && n > i && ((nw[i] >> (t[i] & 31u)) & 1u);
                    if (alive) {
                        if ((s_acc[t[i] >> 5] >> (t[i] & 31u)) & 1u) hit = true;
                        held[p] = static_cast<uint16_t>(t[i]);
                        ++p;
                    }
                }
            }
            // the list's entries join the held set
            bool ovf = false;
            if (act && (h >> 31)) hit = true;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                if (__ballot(cnt > static_cast<uint32_t>(m)) != 0 && cnt > static_cast<uint32_t>(m)) {
                    const uint32_t ew = m < 2 ? e4.x : e4.y;
                    const uint32_t sl = cnt == 1 ? (h & 0xffffu) : ((ew >> (16 * (m & 1))) & 0xffffu);
                    if (p < kTriHeld) {
                        held[p] = static_cast<uint16_t>(sl);
                        ++p;
                    } else {
                        ovf = true;
                    }
                }
            }
            if (__ballot(cnt > 4u) != 0 && cnt > 4u) {
                for (uint32_t m = 4; m < cnt; ++m) {
                    const uint32_t sl = tri.slot[(h & 0xffffffu) + m];
                    if (p < kTriHeld) {
                        held[p] = static_cast<uint16_t>(sl);
                        ++p;
                    } else {
                        ovf = true;
                    }
                }
            }
            // depth 1 (a first position accepts on b) and depth 2 (the pair list accepts)
            if (act && (s_flags[b] & 2u)) hit = true;
            if (act && hk >= 1 && ((s_pacc[(yb << 3) | (b >> 5)] >> (b & 31u)) & 1u)) hit = true;
            if (act) {
                n = p;
                if (ovf) {
                    found = -1;
                    stop = true;
                } else if (hit && ga + k >= it.cs) {
                    found = static_cast<int32_t>(ga + k - it.ib + 1);
                    stop = true;
                }
            }
        }
        // ---- group end: carry the last two bytes, move on, finish items
        if (running && !stop) {
            const int64_t gend = ga + 16;
            if (!more) {
                stop = true;  // the chunk's end: no accept at or after its start
            } else {
                const int32_t stepped = 16 - s;
                x = stepped >= 2 ? group_byte(v, 14) : y;
                y = group_byte(v, 15);
                hist = (hist + stepped) < 2 ? hist + stepped : 2;
                q = gend;
                v = vn;
                ++groups;
                // another chunk of this input may have found an earlier match meanwhile
                if ((groups & 15) == 0 && settled(sc.best, it, q)) stop = true;
            }
        }
        if (running && stop) {
            if (found > 0) atomicMin(&sc.best[it.in], found);
            if (found < 0) {
                sc.ovf_items[atomicAdd(&sc.counter[2], 1u)] = static_cast<int32_t>(it.id);
                __atomic_store_n(&sc.deferred[it.in], 1, __ATOMIC_RELAXED);
            }
            running = false;
        }
    }
}

// ------------------------------------------------------------------------- lds tier ----

template <bool kPairs, int kCap>
__global__ __launch_bounds__(64 * kChainWaves) void nfa_chain_lds_kernel(
    NfaChainDev d, const uint8_t *__restrict__ data, const int64_t *__restrict__ off,
    int32_t count, ChainScratch sc, int64_t chunk, bool from_list, int32_t *ovf_out) {
    __shared__ int32_t lists[kChainWaves][2][kCap];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const size_t plane64 = static_cast<size_t>(d.tiles + 1) * kTileWords;  // words per byte plane
    const int32_t maxk = *sc.maxk;
    // every item, or only those the register tier could not finish
    const int64_t items = from_list ? static_cast<int64_t>(sc.counter[2]) : sc.rbase[maxk];
    // An empty list is the common case after the register or lane tier: leave before the
    // hand-out, whose one atomic per wave on one counter would serialize the whole grid.
    if (items == 0) return;
    const int32_t *list = from_list ? sc.ovf_items : nullptr;
    const int64_t data_end = off[count];
    const int64_t overlap = d.max_chain > 0 ? d.max_chain - 1 : 0;
    Item it;
    ItemPool pool;
    const uint32_t per = pool_block(items, static_cast<int64_t>(gridDim.x) * kChainWaves);
    while (next_item(&sc.counter[1], per, pool, list, sc, off, maxk, items, chunk, overlap,
                     lane, it)) {
        // wave-uniform; an input already handed to the tile tier is rerun there whole
        bool stop = settled(sc.best, it, it.cs) ||
                    __atomic_load_n(&ovf_out[it.in], __ATOMIC_RELAXED) != 0;
        int32_t *A = lists[wv][0];
        int32_t *B = lists[wv][1];
        int32_t k = 0;
        bool s0 = true, ovf = false;
        int32_t d1 = 256;  // pair mode: byte of the depth-1 set; 256 = none
        int32_t found = 0;
        const int64_t abase =
            it.ps - static_cast<int64_t>((reinterpret_cast<uintptr_t>(data) + it.ps) & 3u);
        for (int64_t w0 = abase; w0 < it.ce && !stop; w0 += 256) {
            const uint32_t word = chain_load_word(data, w0 + 4 * lane, it.ce, data_end);
            const int32_t lo = it.ps > w0 ? static_cast<int32_t>(it.ps - w0) : 0;
            const int32_t hi = (it.ce - w0) < 256 ? static_cast<int32_t>(it.ce - w0) : 256;
            const int32_t hit_from = it.cs > w0 ? static_cast<int32_t>(it.cs - w0) : 0;
            for (int32_t r = lo; r < hi; ++r) {
                const uint32_t b = window_byte(word, r);
                const uint32_t fl = d.bflags[b];
                int32_t add_off = 0, add_cnt = 0;
                bool add_acc = false;
                if (kPairs) {
                    const int32_t *pe = d.pair + 2 * (d1 * 256 + static_cast<int32_t>(b));
                    add_off = pe[0];
                    add_cnt = pe[1] & INT_MAX;
                    add_acc = pe[1] < 0;
                } else if (s0) {
                    add_off = d.first_off[b];
                    add_cnt = d.first_off[b + 1] - add_off;
                }
                const int32_t *add_src = kPairs ? d.pair_slot : d.first_slot;
                int32_t nk = 0;
                bool hit = add_acc || (s0 && (fl & 2u));
                for (int32_t j0 = 0; j0 < k; j0 += 64) {
                    const int32_t jj = j0 + lane;
                    bool alive = false;
                    int32_t t = 0;
                    if (jj < k) {
                        t = A[jj] + 1;
                        const uint32_t tile = static_cast<uint32_t>(t) >> 12;
                        const uint32_t sl = static_cast<uint32_t>(t) & 4095u;
                        const uint64_t nw =
                            d.next[static_cast<size_t>(b) * plane64 + (static_cast<uint32_t>(t) >> 6)];
                        const uint64_t aw = d.acc[tile * kTileWords + (sl >> 6)];
                        alive = (nw >> (sl & 63)) & 1ull;
                        hit |= alive && ((aw >> (sl & 63)) & 1ull);
                    }
                    const uint64_t m = __ballot(alive);
                    const int32_t cnt = __popcll(m);
                    if (nk + cnt > kCap) {
                        ovf = true;
                        break;
                    }
                    if (alive) B[nk + lane_rank(m)] = t;
                    nk += cnt;
                }
                if (!ovf && add_cnt > 0) {
                    if (nk + add_cnt > kCap) {
                        ovf = true;
                    } else {
                        for (int32_t jj = lane; jj < add_cnt; jj += 64)
                            B[nk + jj] = add_src[add_off + jj];
                        nk += add_cnt;
                    }
                }
                if (ovf) {
                    stop = true;
                    break;
                }
                if (kPairs) d1 = (s0 && (fl & 1u)) ? static_cast<int32_t>(b) : 256;
                if (!d.start_always) s0 = s0 && (fl & 4u);
                wave_sync();
                int32_t *tmp = A;
                A = B;
                B = tmp;
                k = nk;
                if (__ballot(hit) != 0 && r >= hit_from) {
                    found = static_cast<int32_t>(w0 + r - it.ib + 1);
                    stop = true;
                    break;
                }
                if (k == 0 && !s0 && (!kPairs || d1 == 256)) {
                    stop = true;
                    break;
                }
            }
            // another chunk of this input may have found an earlier match meanwhile
            if (!stop && settled(sc.best, it, w0 + 256)) stop = true;
        }
        if (lane == 0) {
            if (found > 0) atomicMin(&sc.best[it.in], found);
            if (ovf) {
                ovf_out[it.in] = 1;
                atomicAdd(&sc.counter[3], 1u);  // the tile tier runs only when this is nonzero
            }
        }
    }
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

int32_t nfa_chain_chunk_bytes(int32_t requested, int64_t total_bytes) {
    int64_t c = requested > 0 ? requested : env_int("GFSM_CHAIN_CHUNK", 0);
    if (c <= 0) {
        // 4 KiB items for large batches; a small batch (a few requests of a serving launch)
        // is cut finer, down to 512 B, so that it still makes about 4096 items: a request's
        // latency is the walk of its longest item, about 200 ns per byte
        c = 4096;
        while (c > 512 && total_bytes / c < 4096) c >>= 1;
    }
    return static_cast<int32_t>(c < 64 ? 64 : (c > (1 << 30) ? (1 << 30) : c));
}

// Chunk length actually used: inputs are cut only when the start state is always active.
static int64_t effective_chunk(const NfaChainDev &d, int32_t chunk_bytes) {
    return d.start_always ? static_cast<int64_t>(chunk_bytes) : (int64_t(1) << 62);
}

// Most items a launch can have: every input contributes ceil(length / chunk) of them.
static size_t item_capacity(const NfaChainDev &d, int32_t count, int64_t total_bytes,
                            int32_t chunk_bytes) {
    return static_cast<size_t>(total_bytes / effective_chunk(d, chunk_bytes)) +
           static_cast<size_t>(count) + 1;
}

// Capacity of the per-round arrays: more than the most chunks any input can have.
static int32_t round_capacity(int64_t total_bytes, int64_t chunk) {
    const int64_t k = total_bytes / chunk + 3;
    return static_cast<int32_t>(k < (int64_t(1) << 30) ? k : (int64_t(1) << 30));
}

size_t nfa_chain_scratch_bytes(const NfaChainDev &d, int32_t count, int64_t total_bytes,
                               int32_t chunk_bytes) {
    const size_t c = static_cast<size_t>(count);
    const size_t K = static_cast<size_t>(round_capacity(total_bytes, effective_chunk(d, chunk_bytes)));
    return 5 * align256(c * 4) + align256(item_capacity(d, count, total_bytes, chunk_bytes) * 4) +
           3 * align256(K * 4) + align256(K * 8) + 256 + 256;
}

void launch_nfa_chain(const NfaChainDev &d, const NfaTileDev &tiles, const uint8_t *data,
                      const int64_t *off, int32_t count, int64_t total_bytes, int32_t *flags,
                      int32_t *lens, void *scratch, int32_t chunk_bytes, bool pairs,
                      bool paired, const NfaTriDev *tri, hipStream_t stream) {
    if (count <= 0) return;
    const int64_t chunk = effective_chunk(d, chunk_bytes);
    // item numbers are 32-bit
    if (item_capacity(d, count, total_bytes, chunk_bytes) >= (size_t(1) << 31))
        throw Error(1, "chain: more than 2^31 (input, chunk) items; raise chunk_bytes");
    const size_t c = static_cast<size_t>(count);
    ChainScratch sc;
    sc.K = round_capacity(total_bytes, chunk);
    const size_t K = static_cast<size_t>(sc.K);
    char *p = static_cast<char *>(scratch);
    auto take = [&](size_t bytes) {
        char *q = p;
        p += align256(bytes);
        return q;
    };
    sc.best = reinterpret_cast<int32_t *>(take(c * 4));
    sc.ovf_items = reinterpret_cast<int32_t *>(
        take(item_capacity(d, count, total_bytes, chunk_bytes) * 4));
    sc.ovf2 = reinterpret_cast<int32_t *>(take(c * 4));
    sc.deferred = reinterpret_cast<int32_t *>(take(c * 4));
    sc.nch = reinterpret_cast<int32_t *>(take(c * 4));
    sc.perm = reinterpret_cast<int32_t *>(take(c * 4));
    sc.hist = reinterpret_cast<int32_t *>(take(K * 4));
    sc.cursor = reinterpret_cast<int32_t *>(take(K * 4));
    sc.rcount = reinterpret_cast<int32_t *>(take(K * 4));
    sc.rbase = reinterpret_cast<int64_t *>(take(K * 8));
    sc.maxk = reinterpret_cast<int32_t *>(take(4));
    sc.counter = reinterpret_cast<unsigned int *>(take(16));
    chain_zero_kernel<<<(sc.K + 255) / 256, 256, 0, stream>>>(sc);
    GFSM_HIP(hipGetLastError());
    chain_prep_kernel<<<(count + 255) / 256, 256, 0, stream>>>(off, count, chunk, sc);
    GFSM_HIP(hipGetLastError());
    chain_rounds_kernel<<<1, 1024, 0, stream>>>(sc);
    GFSM_HIP(hipGetLastError());
    chain_scatter_kernel<<<(count + 255) / 256, 256, 0, stream>>>(sc, count);
    GFSM_HIP(hipGetLastError());
    const int threads = 64 * kChainWaves;
    // persistent grids: as many blocks as stay resident at once on every CU
    const int grid_reg = persistent_blocks(
        reinterpret_cast<const void *>(&nfa_chain_reg_kernel<true, true>), threads);
    const int grid_pair = persistent_blocks(
        reinterpret_cast<const void *>(&nfa_chain_pair_kernel<true, true>), threads);
    const int grid_lds_pairs = persistent_blocks(
        reinterpret_cast<const void *>(&nfa_chain_lds_kernel<true, 1024>), threads);
    const int grid_lds_first = persistent_blocks(
        reinterpret_cast<const void *>(&nfa_chain_lds_kernel<false, 1024>), threads);
    if (d.pair && pairs) {
        const bool small = d.tiles <= kSmallTiles && d.pair_slot16 != nullptr;
        if (tri && d.start_always && small) {
            // GFSM_CHAIN_TRI_LANES caps the lanes (sweeps): fewer lanes than items leave the
            // later chunks of an input to start after its earlier ones, which may settle it
            int blocks = persistent_blocks(reinterpret_cast<const void *>(&nfa_chain_tri_kernel),
                                           64 * kTriWaves);
            const int64_t cap = (env_int("GFSM_CHAIN_TRI_LANES", 0) + 64 * kTriWaves - 1) / (64 * kTriWaves);
            if (cap > 0 && cap < blocks) blocks = static_cast<int>(cap);
            nfa_chain_tri_kernel<<<blocks, 64 * kTriWaves, 0, stream>>>(d, *tri, data, off, count,
                                                                         sc, chunk);
        }
        else if (paired && d.start_always && small)
            nfa_chain_pair_kernel<true, true><<<grid_pair, threads, 0, stream>>>(d, data, off, count, sc, chunk);
        else if (paired && d.start_always)
            nfa_chain_pair_kernel<true, false><<<grid_pair, threads, 0, stream>>>(d, data, off, count, sc, chunk);
        else if (paired && small)
            nfa_chain_pair_kernel<false, true><<<grid_pair, threads, 0, stream>>>(d, data, off, count, sc, chunk);
        else if (paired)
            nfa_chain_pair_kernel<false, false><<<grid_pair, threads, 0, stream>>>(d, data, off, count, sc, chunk);
        else if (d.start_always && small)
            nfa_chain_reg_kernel<true, true><<<grid_reg, threads, 0, stream>>>(d, data, off, count, sc, chunk);
        else if (d.start_always)
            nfa_chain_reg_kernel<true, false><<<grid_reg, threads, 0, stream>>>(d, data, off, count, sc, chunk);
        else if (small)
            nfa_chain_reg_kernel<false, true><<<grid_reg, threads, 0, stream>>>(d, data, off, count, sc, chunk);
        else
            nfa_chain_reg_kernel<false, false><<<grid_reg, threads, 0, stream>>>(d, data, off, count, sc, chunk);
        GFSM_HIP(hipGetLastError());
        nfa_chain_lds_kernel<true, 1024><<<grid_lds_pairs, threads, 0, stream>>>(
            d, data, off, count, sc, chunk, true, sc.ovf2);
    } else {
        nfa_chain_lds_kernel<false, 1024><<<grid_lds_first, threads, 0, stream>>>(
            d, data, off, count, sc, chunk, false, sc.ovf2);
    }
    GFSM_HIP(hipGetLastError());
    launch_nfa_bitpacked_core(tiles, data, off, count, sc.best, sc.ovf2, &sc.counter[3], stream);
    launch_nfa_best_finish(sc.best, count, flags, lens, stream);
}

}  // namespace gfsm
