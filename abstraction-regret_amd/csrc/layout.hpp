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
    std::vector<uint64_t> mask, first, hassucc, acc;  // see NfaTileDev
    std::vector<int32_t> slot_of_state;               // renumbering (-1 for the start)
};

// Validates the CSR (throws Error(GFSM_EINVAL) on a malformed automaton).
void validate_nfa(const NfaInput &a);
WorklistHost build_worklist(const NfaInput &a);
TileHost build_tiles(const NfaInput &a);

// DFA: entries = successor | accept(successor) << (15 or 31).
struct DfaHost {
    std::vector<uint16_t> t16;
    std::vector<uint32_t> t32;
    bool wide = false;
    bool start_accepts = false;
};
DfaHost build_dfa(int32_t n, int32_t start, const uint8_t *accept, const int32_t *trans);

}  // namespace gfsm
