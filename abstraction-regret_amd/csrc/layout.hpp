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

// DFA device layout. States are renumbered so that the accepting ones come last; the
// table is stored by byte (column c holds the successor of every state on byte c) and each
// entry is the successor's byte offset inside a column (state * entry size). A step is then
// one multiply-add and one load, and "the successor accepts" is "entry >= accept_from".
struct DfaHost {
    std::vector<uint16_t> t16;          // [256][n] when n <= 32767
    std::vector<uint32_t> t32;          // [256][n] otherwise
    bool wide = false;
    bool start_accepts = false;
    uint32_t start_entry = 0;           // renumbered start state, scaled
    uint32_t accept_from = 0;           // scaled id of the first accepting state
    uint32_t col_bytes = 0;             // n * entry size
};
DfaHost build_dfa(int32_t n, int32_t start, const uint8_t *accept, const int32_t *trans);

}  // namespace gfsm
