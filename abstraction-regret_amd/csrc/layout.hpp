// Host-side preprocessing of an upstream CSR NFA / dense DFA into the device layouts.
//
// Pure C++ (no HIP calls): the C ABI uploads what these functions produce. Input layouts
// follow `src/gpufsm/core/nfa.py:43-65` and `src/gpufsm/core/dfa.py:25-37`.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace gfsm {

struct NfaInput {
    int32_t n, start;
    const uint8_t *accept;
    const int32_t *srp, *st, *ss, *erp, *et;
};

struct WorklistHost {
    std::vector<int32_t> row, bigidx, bigoff, erow, etgt, start_list;
    std::vector<uint32_t> edge;
    std::vector<uint64_t> accw;
    int32_t nwords = 0;
    bool start_accepts = false;
    bool has_eps = false;
};

struct TileHost {
    bool ok = false;
    std::string why;                    // reason when !ok
    int32_t tiles = 0;
    int32_t max_chain = 0;
    bool start_always = false;
    uint64_t selfmask[4] = {0, 0, 0, 0};
    std::vector<uint64_t> mask, first, hassucc, acc;  // see NfaTileDev; acc has one extra
                                                      // all-zero tile
    std::vector<int32_t> slot_of_state;               // renumbering (-1 for the start)
    // successors of the start state by byte: slots of first positions whose class holds c
    std::vector<int32_t> first_off;                   // [257]
    std::vector<int32_t> first_slot;                  // [first_off[256]]
    std::vector<uint32_t> first_accepts;              // [256] some first position accepts
    // chain continuation by byte, [256][tiles + 1][64] (byte-major: one plane of
    // (tiles + 1) * 64 words per byte, so slot t's word on byte c is plane c, word t >> 6):
    // bit t of next[c][tile] is set when slot t - 1 has a chain successor and the class of
    // slot t holds c (mask & hassucc << 1). The extra all-zero tile lets a step read slot
    // last + 1 without a bounds test; bit 0 of every tile is zero (chains never cross
    // tiles), so a lookup of slot 0 never continues anything.
    std::vector<uint64_t> next;
    // the same by bits: the start state has a successor on c / one that accepts
    uint64_t first_any[4] = {0, 0, 0, 0};
    uint64_t first_acc[4] = {0, 0, 0, 0};
    // second positions by byte pair: pair[c1 * 256 + c2] = {offset, count | accepts << 31}
    // into pair_slot, the slots v = succ(u) of first positions u whose start class holds c1
    // and whose own class holds c2. Empty when the lists would exceed kMaxPairSlots.
    std::vector<int32_t> pair;                        // [257 * 256 * 2] or empty; row 256
                                                      // (no depth-1 set) is all zero
    std::vector<int32_t> pair_slot;                   // + 64 zero padding entries
    std::vector<uint32_t> bflags;                     // [256] first_any | first_acc << 1 |
                                                      // start self-loop << 2
};

// Validates the CSR (throws Error(GFSM_EINVAL) on a malformed automaton).
void validate_nfa(const NfaInput &a);
WorklistHost build_worklist(const NfaInput &a);
TileHost build_tiles(const NfaInput &a);

// Third positions by byte triple, for chain automata whose start state loops on every byte:
// tri(c1, c2, c3) = { v + 1 : v in pair(c1, c2), bit v + 1 of next[c3] }, the slots entered
// at depth 3 on c3 after c1 c2 (the depth-1 and depth-2 sets depend on the last two bytes
// alone when the start state is always active). One 32-bit header per triple, at index
// c1 << 16 | c2 << 8 | c3: accepts << 31 | count << 24 | low 24 bits, which hold the slot
// itself when count == 1 and otherwise the list's offset in `slot`, a multiple of 4 (the
// first four entries are one 8-byte load). Built only when every slot is below 65536, no
// list exceeds 127 entries and the lists take at most kMaxTripleSlots entries.
struct TripleHost {
    bool ok = false;
    std::string why;                    // reason when !ok
    std::vector<uint32_t> hdr;          // [1 << 24]
    std::vector<uint32_t> nonempty;     // [1 << 19]: bit i set when hdr[i] has entries (2 MiB,
                                        // small enough to stay in L2 while hdr does not)
    std::vector<uint16_t> slot;         // lists of two or more entries, + 8 zero entries
    std::vector<uint32_t> pair_acc;     // [2048]: bit c1 << 8 | c2 set when pair(c1, c2)
                                        // holds an accepting slot
    int64_t entries = 0;                // entries over all lists (single ones included)
};
TripleHost build_triples(const TileHost &t);

// The same automaton with its epsilon closure folded into the symbol edges: an edge
// (s, c, t) becomes (s, c, u) for every u in closure(t), duplicates dropped, and the start
// set is closure(start). Simulating it without any epsilon pass gives the verdicts of
// `reference.py:36-65` at every byte: closure(union of targets) is the union of the
// targets' closures, and a successor set is empty exactly when its closure is.
struct FoldedHost {
    bool ok = false;                    // false: the folded edges would exceed the cap
    std::string why;
    std::vector<int32_t> srp, st, ss;   // CSR of the folded symbol edges (no epsilon edges)
    std::vector<int32_t> start_set;     // closure(start), start first
    bool start_accepts = false;
};
// cap: most folded edges accepted (0 = max(4 * nnz, 2^22)).
FoldedHost fold_epsilon(const NfaInput &a, int64_t cap = 0);

// Device layout of the general techniques (any NFA), built from the folded automaton.
//  * rows: label-sorted folded rows with the worklist's per-label index (WorklistHost of
//    the folded automaton; start_list = closure(start));
//  * byte-major gather lists (n <= 65536): the folded edges labelled c as
//    (source << 16 | target) at [boff[c], boff[c + 1]), the ANY edges at
//    [boff[256], boff[257]);
//  * start_bits: closure(start) as a bitset.
//  * masks that take edges out of the gather lists (n <= 65536, see DiagMode):
//    - first masks fmask[c][nwords], when the start state stays active on every byte (a
//      folded self-loop on each byte or ANY): the folded successors of the start on c or
//      ANY, OR'ed into every step; the start's edges leave the lists;
//    - diagonal masks dmask[c][k][nwords] for up to kMaxDiag offsets doff[k] (1 <= |d| <= 63):
//      bit t set when t - d -c-> t or t - d -ANY-> t (indexed by target); a step adds
//      (set shifted by d across words) & mask, and those edges leave the lists.
//    The lists keep every other edge, so lists and masks together are the folded relation.
constexpr int kMaxDiag = 4;  // also the size of NfaGeneralDev::doff
enum class DiagMode { off, automatic, force };
struct GeneralHost {
    bool ok = false;                    // folded (and the rows built)
    std::string why;
    WorklistHost rows;
    bool gather_ok = false;             // n <= 65536: the byte-major lists exist
    std::vector<int32_t> boff;          // [258]
    std::vector<uint32_t> bedge;
    std::vector<uint64_t> start_bits;   // [nwords]
    std::vector<uint64_t> fmask;        // [256][nwords] or empty
    std::vector<uint64_t> dmask;        // [256][ndiag][nwords] or empty
    int32_t ndiag = 0;
    int32_t doff[kMaxDiag] = {0, 0, 0, 0};
};
// mode automatic: a mask is kept when it replaces at least one list edge per 64 states per
// byte on average (one 8-byte mask word read instead of >= 1 four-byte edge per word);
// force: every start loop and the kMaxDiag most used offsets (tests).
GeneralHost build_general(const NfaInput &a, DiagMode mode = DiagMode::automatic);

// DFA device layout. States are renumbered so that the accepting ones come last; the
// table is stored by byte (column c holds the successor of every state on byte c).
//  * Up to 256 states: byte entries, the successor's id, in columns of 256 (padded), so the
//    successor of entry e on byte c is at (c << 8) | e: 64 KiB.
//  * Up to 32767 states: 16-bit entries, above that 32-bit, each the successor's byte
//    offset inside its column (state * entry size), so a step is one multiply-add and one
//    load.
// Either way "the successor accepts" is "entry >= accept_from".
struct DfaHost {
    std::vector<uint8_t> t8;            // [256][256] when n <= 256
    std::vector<uint16_t> t16;          // [256][n] when 256 < n <= 32767
    std::vector<uint32_t> t32;          // [256][n] otherwise
    uint32_t esz = 2;                   // entry size in bytes: 1, 2 or 4
    bool wide = false;                  // esz == 4
    bool start_accepts = false;
    uint32_t start_entry = 0;           // renumbered start state, scaled
    uint32_t accept_from = 0;           // scaled id of the first accepting state
    uint32_t col_bytes = 0;             // bytes per column: 256 (byte entries) or n * esz
};
DfaHost build_dfa(int32_t n, int32_t start, const uint8_t *accept, const int32_t *trans);

}  // namespace gfsm
