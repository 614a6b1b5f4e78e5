// Internal declarations shared by the translation units of libgpufsm_hip.so.
//
// Layout of the device-resident automata, the error plumbing used by the C ABI, and the
// host-side launchers each kernel translation unit exports (kernels are launched from the
// translation unit that defines them; no relocatable device code).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace gfsm {

constexpr int kAnySymbol = 256;          // wildcard label (upstream ANY_SYMBOL)
constexpr int kAlphabet = 256;           // DFA row stride (upstream ALPHABET)
constexpr int kLabelBits = 9;            // packed edge = target << 9 | label (label <= 256)
constexpr int kMaxPackedStates = 1 << 23;
constexpr int kBigRow = 16;              // rows longer than this get a per-label index
constexpr int kBigStride = 258;          // per-label index: 257 label starts + row end
constexpr int kTileStates = 4096;        // bitpacked tile: 64 lanes x 64-bit words
constexpr int kTileWords = 64;
constexpr int64_t kMaxPairSlots = int64_t(1) << 24;  // byte-pair list cap (entries)
// byte-triple lists: entries over all lists, and padded entries (offsets are 24-bit)
constexpr int64_t kMaxTripleSlots = int64_t(1) << 24;

// Thrown inside the library, converted to a status code at the C boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Records `msg` as the calling thread's last error (gfsm_last_error) and returns `code`: the
// conversion every C entry point applies to a failure.
int report_error(int code, const std::string &msg);

#define GFSM_HIP(call)                                                                    \
    do {                                                                                  \
        hipError_t err_ = (call);                                                         \
        if (err_ != hipSuccess)                                                           \
            throw ::gfsm::Error(4, std::string(#call) + " failed at " + __FILE__ + ":" +  \
                                       std::to_string(__LINE__) + ": " +                  \
                                       hipGetErrorString(err_));                          \
    } while (0)

// ----- NFA ------------------------------------------------------------------------------

// Worklist layout: label-sorted CSR rows with packed edges, and a per-label index for
// the few states whose rows are long (e.g. the start state of a pattern union).
struct NfaWorklistDev {
    const int32_t *row;        // [n + 1]
    const uint32_t *edge;      // [nnz] target << 9 | label, sorted by label in each row
    const int32_t *bigidx;     // [n] index into bigoff, or -1
    const int32_t *bigoff;     // [nbig * 258] absolute edge offset of the first label >= c
    const int32_t *erow;       // [n + 1]
    const int32_t *etgt;       // [nnz_eps]
    const uint64_t *accw;      // [nwords] accept bitset
    const int32_t *start_list; // [k0] start closure
    int32_t n, nwords, k0, has_eps;
};

// General layout (any NFA, epsilon closure folded on the host; see layout.hpp): label-sorted
// folded rows with the worklist's per-label index, the closure of the start state, and the
// byte-major gather lists of the bit-parallel step (null above 65536 states).
struct NfaGeneralDev {
    const int32_t *row;        // [n + 1]
    const uint32_t *edge;      // [nnz'] target << 9 | label, label-sorted per row
    const int32_t *bigidx;     // [n]
    const int32_t *bigoff;     // [nbig * 258]
    const uint64_t *accw;      // [nwords]
    const int32_t *start_list; // [k0] closure(start)
    const uint64_t *start_bits;// [nwords] closure(start)
    const int32_t *boff;       // [258] or null: edges labelled c at [boff[c], boff[c + 1]),
                               // ANY edges at [boff[256], boff[257])
    const uint32_t *bedge;     // [nedges] source << 16 | target
    const uint64_t *fmask;     // [256][nwords] or null: successors of an always-active start
    const uint64_t *dmask;     // [256][ndiag][nwords]: targets of edges t - doff[k] -> t
    int32_t n, nwords, k0, nedges;
    int32_t ndiag;             // at most 4 (kMaxDiag)
    int32_t doff[4];
};
constexpr int kRegMaxWords = 8;  // register technique: at most 512 states

// Bitpacked chain layout: states renumbered into tiles of 4096 so that every chain is
// contiguous inside one tile and the successor of bit i is bit i + 1.
struct NfaTileDev {
    const uint64_t *mask;      // [tiles][256][64] bit set <=> byte in the state's class
    const uint64_t *first;     // [tiles][64] successors of the start state
    const uint64_t *hassucc;   // [tiles][64] the state has a chain successor (bit i -> i+1)
    const uint64_t *acc;       // [tiles][64] accepting states
    uint64_t selfmask[4];      // bytes on which the start state stays active
    int32_t tiles;
    int32_t start_always;      // self-loop on every byte
};

// Sparse chain layout (same renumbering as the tiles): the active set is a list of slots;
// slot s continues to s + 1 on byte c when bit s + 1 of next[tile][c] is set. The start
// state's successors come either from byte-pair lists of second positions (pair != null) or
// from per-byte lists of first positions (see layout.hpp).
struct NfaChainDev {
    const uint64_t *next;          // [256][tiles + 1][64]: mask & (hassucc << 1), by byte
    const uint64_t *acc;           // [tiles + 1][64] accepting slots (last tile all zero)
    const int32_t *first_off;      // [257]
    const int32_t *first_slot;
    const int32_t *pair;           // [257 * 256][2] {offset, count | accepts << 31}, or
                                   // null; row 256 is all zero (no depth-1 set)
    const int32_t *pair_slot;      // lists + 64 padding entries
    const uint16_t *pair_slot16;   // the same as uint16 when every slot is < 65536, else null
    const uint32_t *bflags;        // [256] bit 0 first_any, 1 first_acc, 2 start self-loop
    uint64_t first_any[4];         // bytes on which the start state has a successor
    uint64_t first_acc[4];         // ... an accepting one
    uint64_t selfmask[4];          // bytes on which the start state stays active
    int32_t tiles;
    int32_t max_chain;
    int32_t start_always;
    int32_t pad;
};

// ----- DFA ------------------------------------------------------------------------------

// Table by byte (see layout.hpp): byte entries (state ids, 256-entry columns) up to 256
// states, else entries = successor's byte offset in a column, so the step is
// `e = *(T *)(table + byte * col_bytes + e)`; e >= accept_from means accepting.
struct DfaDev {
    const void *table;         // uint8 [256][256], uint16 [256][n] or uint32 [256][n]
    int32_t n, wide;           // wide: uint32 entries
    int32_t esz;               // entry size: 1, 2 or 4
    uint32_t start_entry, accept_from, col_bytes;
};

// ----- launchers (asynchronous on `stream`) ---------------------------------------------

void launch_fill_result(int32_t *flags, int32_t *lens, int32_t count, int32_t flag,
                        int32_t len, hipStream_t stream);

void launch_dfa_gather(const DfaDev &d, const uint8_t *data, const int64_t *off, int32_t count,
                       int32_t *flags, int32_t *lens, hipStream_t stream);
void launch_dfa_lds(const DfaDev &d, const uint8_t *data, const int64_t *off, int32_t count,
                    int32_t *flags, int32_t *lens, hipStream_t stream);
// Bytes of the device table (what an LDS walk stages).
size_t dfa_table_bytes(const DfaDev &d);

// Split DFA walk (exact, in-input parallel): geometry and scratch of one launch.
struct DfaSplitPlan {
    int64_t seg_bytes = 0;     // segment length C (multiple of the checkpoint interval)
    int32_t ckpt_shift = 0;    // log2 of the checkpoint interval R
    int64_t segments = 0;      // ceil(total_bytes / C)
    int32_t chains = 1;        // interleaved chains per lane
    bool maps = false;         // unresolved heads get full transition maps (LDS tables)
    int32_t map_cap = 0;       // most heads mapped per launch (the rest are re-walked)
    size_t scratch_bytes = 0;  // device scratch the launch needs
};
// seg_bytes: requested segment length (a power of two in [256, 1 MiB]); 0 = default
DfaSplitPlan dfa_split_plan(const DfaDev &d, int64_t total_bytes, int32_t count,
                            int64_t seg_bytes = 0);
void launch_dfa_split(const DfaDev &d, const DfaSplitPlan &plan, const uint8_t *data,
                      const int64_t *off, int32_t count, int64_t total_bytes, int32_t *flags,
                      int32_t *lens, void *scratch, hipStream_t stream);
// Compute units of the current device.
int32_t device_cu_count();
// Raises a kernel's dynamic-LDS limit on the current device to at least `bytes` (a no-op at
// or below the default 64 KiB). Thread-safe; the attribute is set once per device and size.
void ensure_dynamic_lds(const void *kernel, size_t bytes);
// Blocks a persistent launch of `kernel` with `threads` per block keeps resident on every
// CU of the current device (occupancy x CUs), cached per device and kernel.
int persistent_blocks(const void *kernel, int threads);
// Positive integer from the environment (tuning overrides), else fallback.
int64_t env_int(const char *name, int64_t fallback);

// General techniques (general.hip).
bool nfa_gather_fits(const NfaGeneralDev &d);
void launch_nfa_gather(const NfaGeneralDev &d, const uint8_t *data, const int64_t *off,
                       int32_t count, int32_t *flags, int32_t *lens, hipStream_t stream);
void nfa_wl2_geometry(const NfaGeneralDev &d, int32_t count, int32_t *grid,
                      int64_t *scratch_elems);
void launch_nfa_wl2(const NfaGeneralDev &d, const uint8_t *data, const int64_t *off,
                    int32_t count, int32_t *flags, int32_t *lens, int32_t *scratch,
                    int32_t grid, hipStream_t stream);
void launch_nfa_register(const NfaGeneralDev &d, const uint8_t *data, const int64_t *off,
                         int32_t count, int32_t *flags, int32_t *lens, hipStream_t stream);

void launch_nfa_worklist(const NfaWorklistDev &d, const uint8_t *data, const int64_t *off,
                         int32_t count, int32_t *flags, int32_t *lens, int32_t *scratch,
                         int32_t grid, hipStream_t stream);
// Blocks and global list scratch (int32 elements) the worklist launch needs.
void nfa_worklist_geometry(const NfaWorklistDev &d, int32_t count, int32_t *grid,
                           int64_t *scratch_elems);

void launch_nfa_bitpacked(const NfaTileDev &d, const uint8_t *data, const int64_t *off,
                          int32_t count, int32_t *flags, int32_t *lens, int32_t *best,
                          hipStream_t stream);
// The tiled pass alone over inputs with only[i] != 0 (all when only is null), folding its
// first-accept positions into best[] (INT_MAX = none yet) with atomicMin. only_count (or
// null): the number of such inputs, when it is known on the device only; 0 skips the pass.
void launch_nfa_bitpacked_core(const NfaTileDev &d, const uint8_t *data, const int64_t *off,
                               int32_t count, int32_t *best, const int32_t *only,
                               const unsigned int *only_count, hipStream_t stream);
// best[] (1-based first-accept position, INT_MAX = none) -> (flags, lens).
void launch_nfa_best_finish(const int32_t *best, int32_t count, int32_t *flags, int32_t *lens,
                            hipStream_t stream);

// Sparse chain technique: persistent waves over (input, chunk) work items.
size_t nfa_chain_scratch_bytes(const NfaChainDev &d, int32_t count, int64_t total_bytes,
                               int32_t chunk_bytes);
// chunk_bytes when the caller has no preference: GFSM_CHAIN_CHUNK (read per call), else
// 4096 for batches of 16 MiB or more and down to 512 for smaller ones
int32_t nfa_chain_chunk_bytes(int32_t requested, int64_t total_bytes);
// Byte-triple lists of a chain automaton (TripleHost on the device).
struct NfaTriDev {
    const uint32_t *hdr;       // [1 << 24] headers
    const uint32_t *nonempty;  // [1 << 19] bit i: hdr[i] has entries
    const uint16_t *slot;      // lists of two or more entries
    const uint32_t *pair_acc;  // [2048] accept bits of the byte-pair lists
};
// pairs: use the byte-pair lists (register tier first) when the handle has them.
// paired: the register tier runs two items per wave (32 listed states each) instead of one
// (64 listed states).
// tri: the byte-triple lists (null: none); with them and pairs, the lane tier (one item per
// lane) replaces the register tier.
void launch_nfa_chain(const NfaChainDev &d, const NfaTileDev &tiles, const uint8_t *data,
                      const int64_t *off, int32_t count, int64_t total_bytes, int32_t *flags,
                      int32_t *lens, void *scratch, int32_t chunk_bytes, bool pairs,
                      bool paired, const NfaTriDev *tri, hipStream_t stream);

}  // namespace gfsm
