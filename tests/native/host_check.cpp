// Host-logic check of the device layouts, built with AddressSanitizer and UBSan by
// tests/test_native_host.py (CPU suite; no GPU and no HIP runtime call).
//
// It links the layout builders (csrc/layout.cpp) and the C oracle (oracle/fsm_oracle.c) and
// checks, on random automata:
//   * the epsilon fold: the folded automaton, walked through the layouts the kernels read
//     (label-sorted rows with the per-label index; byte-major gather lists with the start
//     bitset, first masks and diagonal masks, which together must rebuild the folded
//     relation byte by byte), gives the verdict of an independent set simulation of reference.py:36-65
//     (closure after every step) and of the C oracle, byte for byte;
//   * the fold cap (a closure too large to fold is refused, not truncated);
//   * the DFA table (renumbered, byte-major, pre-scaled entries) walked as the kernels walk
//     it against the C oracle's simulate_dfa;
//   * the chain tiles: a pattern union is accepted, every state but the start gets one slot,
//     and an automaton with an epsilon edge or a state with two successors is refused;
//   * the chain lane tier's byte-triple lists against a brute-force enumeration, and its walk
//     against the set simulation;
//   * validation: malformed CSR input raises the library's error.
// Exit status 0 and one summary line on success; the first mismatch is printed and exits 1.
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <set>
#include <utility>
#include <vector>

#include "internal.hpp"
#include "layout.hpp"

extern "C" {
int oracle_nfa_batch(int32_t num_states, int32_t start_state, const uint8_t *accept,
                     const int32_t *sym_row_ptr, const int32_t *sym_targets,
                     const int32_t *sym_symbols, const int32_t *eps_row_ptr,
                     const int32_t *eps_targets, const uint8_t *data, const int64_t *offsets,
                     int32_t num_strings, int32_t *flags, int32_t *lens, int64_t *examined,
                     int32_t indexed, int32_t nthreads);
int oracle_dfa_batch(int32_t num_states, int32_t start_state, const uint8_t *accept,
                     const int32_t *trans, const uint8_t *data, const int64_t *offsets,
                     int32_t num_strings, int32_t *flags, int32_t *lens, int64_t *examined,
                     int32_t nthreads);
}

using namespace gfsm;

namespace {

int g_cases = 0;

[[noreturn]] void fail(const char *what, int a = 0, int b = 0) {
    std::fprintf(stderr, "host_check FAILED: %s (%d, %d)\n", what, a, b);
    std::exit(1);
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 1) {}
    uint64_t next() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    }
    int below(int n) { return static_cast<int>(next() % static_cast<uint64_t>(n)); }
    double unit() { return static_cast<double>(next() >> 11) / 9007199254740992.0; }
};

struct Nfa {
    int32_t n = 0, start = 0;
    std::vector<uint8_t> acc;
    std::vector<int32_t> srp, st, ss, erp, et;
    NfaInput in() const {
        return NfaInput{n, start, acc.data(), srp.data(), st.data(), ss.data(), erp.data(),
                        et.data()};
    }
};

Nfa random_nfa(Rng &r, int32_t n, int alphabet, int max_sym, int max_eps, double p_any,
               double p_acc) {
    Nfa a;
    a.n = n;
    a.start = r.below(n);
    a.acc.resize(n);
    a.srp.push_back(0);
    a.erp.push_back(0);
    for (int32_t s = 0; s < n; ++s) {
        a.acc[s] = r.unit() < p_acc;
        const int ks = r.below(max_sym + 1);
        for (int k = 0; k < ks; ++k) {
            a.st.push_back(r.below(n));
            a.ss.push_back(r.unit() < p_any ? kAnySymbol : r.below(alphabet));
        }
        a.srp.push_back(static_cast<int32_t>(a.st.size()));
        const int ke = r.below(max_eps + 1);
        for (int k = 0; k < ke; ++k) a.et.push_back(r.below(n));
        a.erp.push_back(static_cast<int32_t>(a.et.size()));
    }
    a.st.push_back(0);  // keep data() non-null for empty edge lists
    a.ss.push_back(0);
    a.et.push_back(0);
    a.st.pop_back();
    a.ss.pop_back();
    a.et.pop_back();
    return a;
}

// A pattern union with optional positions: a start state that loops on every byte (ANY, or
// one edge per byte), chains of class-labelled states with epsilon edges skipping optional
// positions, the last state of each chain accepting. reversed: each chain's states are
// numbered downwards, so successor edges run to s - 1 (negative diagonals).
Nfa union_nfa(Rng &r, int patterns, int min_len, int max_len, int width, int alphabet,
              double p_opt, bool any_loop, bool reversed) {
    std::vector<std::vector<std::pair<int32_t, int32_t>>> sym(1);  // (label, target) per state
    std::vector<std::vector<int32_t>> eps(1);
    std::vector<uint8_t> acc(1, 0);
    if (any_loop) sym[0].push_back({kAnySymbol, 0});
    else
        for (int c = 0; c < kAlphabet; ++c) sym[0].push_back({c, 0});
    for (int p = 0; p < patterns; ++p) {
        const int len = min_len + r.below(max_len - min_len + 1);
        const int32_t base = static_cast<int32_t>(sym.size());
        auto id = [&](int j) { return reversed ? base + len - 1 - j : base + j; };
        sym.resize(sym.size() + static_cast<size_t>(len));
        eps.resize(sym.size());
        acc.resize(sym.size(), 0);
        for (int j = 0; j < len; ++j) {
            const int w = 1 + r.below(width);
            const int32_t from = j == 0 ? 0 : id(j - 1);
            for (int q = 0; q < w; ++q) sym[static_cast<size_t>(from)].push_back({r.below(alphabet), id(j)});
            if (j + 1 < len && r.unit() < p_opt) eps[static_cast<size_t>(id(j))].push_back(id(j + 1));
        }
        acc[static_cast<size_t>(id(len - 1))] = 1;
    }
    Nfa a;
    a.n = static_cast<int32_t>(sym.size());
    a.start = 0;
    a.acc = acc;
    a.srp.push_back(0);
    a.erp.push_back(0);
    for (int32_t s = 0; s < a.n; ++s) {
        for (auto [lab, t] : sym[static_cast<size_t>(s)]) {
            a.ss.push_back(lab);
            a.st.push_back(t);
        }
        a.srp.push_back(static_cast<int32_t>(a.st.size()));
        for (int32_t t : eps[static_cast<size_t>(s)]) a.et.push_back(t);
        a.erp.push_back(static_cast<int32_t>(a.et.size()));
    }
    return a;
}

// reference.py:36-65, restated independently of the oracle: sets, closure after each step.
std::set<int32_t> closure(const Nfa &a, std::set<int32_t> s) {
    std::vector<int32_t> stack(s.begin(), s.end());
    while (!stack.empty()) {
        const int32_t x = stack.back();
        stack.pop_back();
        for (int32_t k = a.erp[x]; k < a.erp[x + 1]; ++k)
            if (s.insert(a.et[k]).second) stack.push_back(a.et[k]);
    }
    return s;
}

std::pair<int, int> sim_ref(const Nfa &a, const uint8_t *p, int len) {
    std::set<int32_t> cur = closure(a, {a.start});
    for (int32_t x : cur)
        if (a.acc[x]) return {1, 0};
    for (int i = 0; i < len; ++i) {
        std::set<int32_t> nxt;
        for (int32_t x : cur)
            for (int32_t k = a.srp[x]; k < a.srp[x + 1]; ++k)
                if (a.ss[k] == p[i] || a.ss[k] == kAnySymbol) nxt.insert(a.st[k]);
        if (nxt.empty()) return {0, 0};
        cur = closure(a, nxt);
        for (int32_t x : cur)
            if (a.acc[x]) return {1, i + 1};
    }
    return {0, 0};
}

// The folded rows as the worklist and register kernels read them (for_edges in general.hip).
std::pair<int, int> sim_rows(const WorklistHost &r, const uint8_t *acc, const uint8_t *p,
                             int len) {
    std::vector<uint8_t> in(r.row.size() - 1, 0);
    std::vector<int32_t> cur(r.start_list), nxt;
    for (int32_t x : cur)
        if (acc[x]) return {1, 0};
    for (int i = 0; i < len; ++i) {
        const uint32_t b = p[i];
        nxt.clear();
        auto add = [&](uint32_t t) {
            if (!in[t]) {
                in[t] = 1;
                nxt.push_back(static_cast<int32_t>(t));
            }
        };
        for (int32_t s : cur) {
            const int32_t bi = r.bigidx[s];
            if (bi >= 0) {
                const int32_t *bo = r.bigoff.data() + static_cast<int64_t>(bi) * kBigStride;
                for (int32_t e = bo[b]; e < bo[b + 1]; ++e) add(r.edge[e] >> kLabelBits);
                for (int32_t e = bo[kAnySymbol]; e < bo[kAnySymbol + 1]; ++e)
                    add(r.edge[e] >> kLabelBits);
            } else {
                for (int32_t e = r.row[s]; e < r.row[s + 1]; ++e) {
                    const uint32_t lab = r.edge[e] & ((1u << kLabelBits) - 1u);
                    if (lab == b || lab == static_cast<uint32_t>(kAnySymbol))
                        add(r.edge[e] >> kLabelBits);
                }
            }
        }
        for (int32_t t : nxt) in[t] = 0;
        if (nxt.empty()) return {0, 0};
        cur.swap(nxt);
        for (int32_t x : cur)
            if (acc[x]) return {1, i + 1};
    }
    return {0, 0};
}

// The masks' part of the next set on byte c (gather_masks_into in general.hip).
void mask_step(const GeneralHost &g, const std::vector<uint64_t> &cur, std::vector<uint64_t> &nxt,
               int c) {
    const size_t nw = cur.size();
    const size_t nd = static_cast<size_t>(g.ndiag);
    for (size_t w = 0; w < nw; ++w) {
        uint64_t x = g.fmask.empty() ? 0ull : g.fmask[static_cast<size_t>(c) * nw + w];
        for (size_t k = 0; k < nd; ++k) {
            const uint64_t *m = g.dmask.data() + (static_cast<size_t>(c) * nd + k) * nw;
            const int d = g.doff[k];
            uint64_t sh;  // word w of the set shifted by d
            if (d > 0) {
                sh = cur[w] << d;
                if (w > 0) sh |= cur[w - 1] >> (64 - d);
            } else {
                sh = cur[w] >> -d;
                if (w + 1 < nw) sh |= cur[w + 1] << (64 + d);
            }
            x |= sh & m[w];
        }
        nxt[w] = x;
    }
}

// The byte-major gather lists, masks and start bitset as the bit-parallel kernel reads them.
std::pair<int, int> sim_gather(const GeneralHost &g, const uint8_t *p, int len) {
    const WorklistHost &r = g.rows;
    std::vector<uint64_t> cur(g.start_bits), nxt(cur.size());
    for (size_t w = 0; w < cur.size(); ++w)
        if (cur[w] & r.accw[w]) return {1, 0};
    for (int i = 0; i < len; ++i) {
        const int c = p[i];
        mask_step(g, cur, nxt, c);
        for (int pass = 0; pass < 2; ++pass) {
            const int lab = pass ? kAnySymbol : c;
            for (int32_t e = g.boff[lab]; e < g.boff[lab + 1]; ++e) {
                const uint32_t s = g.bedge[e] >> 16, t = g.bedge[e] & 0xffffu;
                if ((cur[s >> 6] >> (s & 63)) & 1ull) nxt[t >> 6] |= 1ull << (t & 63);
            }
        }
        uint64_t any = 0, hit = 0;
        for (size_t w = 0; w < nxt.size(); ++w) {
            any |= nxt[w];
            hit |= nxt[w] & r.accw[w];
        }
        if (hit) return {1, i + 1};
        if (!any) return {0, 0};
        cur.swap(nxt);
    }
    return {0, 0};
}

// Lists and masks together hold exactly the folded relation on each byte: the pairs
// (s, t) with s -c-> t or s -ANY-> t in the folded rows (so masks hold no pair the rows
// lack), and an always-active start (first masks) does loop on every byte.
void check_gather_partition(const GeneralHost &g, int32_t start, int32_t n) {
    const size_t nw = static_cast<size_t>(g.rows.nwords);
    const size_t nd = static_cast<size_t>(g.ndiag);
    if (g.dmask.size() != nd * nw * kAlphabet) fail("diagonal mask size", n, g.ndiag);
    if (!g.fmask.empty() && g.fmask.size() != nw * kAlphabet) fail("first mask size", n, 0);
    for (size_t k = 0; k < nd; ++k) {
        if (g.doff[k] == 0 || g.doff[k] < -63 || g.doff[k] > 63) fail("diagonal offset", n, g.doff[k]);
        for (size_t q = 0; q < k; ++q)
            if (g.doff[q] == g.doff[k]) fail("repeated diagonal", n, g.doff[k]);
    }
    for (int c = 0; c < kAlphabet; ++c) {
        std::set<std::pair<int32_t, int32_t>> want, got;
        for (int32_t s = 0; s < n; ++s)
            for (int32_t e = g.rows.row[s]; e < g.rows.row[s + 1]; ++e) {
                const uint32_t lab = g.rows.edge[e] & 511u;
                if (lab == static_cast<uint32_t>(c) || lab == kAnySymbol)
                    want.insert({s, static_cast<int32_t>(g.rows.edge[e] >> 9)});
            }
        for (int pass = 0; pass < 2; ++pass) {
            const int lab = pass ? kAnySymbol : c;
            for (int32_t e = g.boff[lab]; e < g.boff[lab + 1]; ++e) {
                const auto pr = std::make_pair(static_cast<int32_t>(g.bedge[e] >> 16),
                                               static_cast<int32_t>(g.bedge[e] & 0xffffu));
                got.insert(pr);  // a pair may sit in both the c and the ANY list
            }
        }
        for (size_t w = 0; w < nw; ++w) {
            const uint64_t fw = g.fmask.empty() ? 0ull : g.fmask[static_cast<size_t>(c) * nw + w];
            for (int b = 0; b < 64; ++b)
                if ((fw >> b) & 1ull) {
                    const int32_t t = static_cast<int32_t>(w * 64) + b;
                    got.insert({start, t});
                }
            for (size_t k = 0; k < nd; ++k) {
                const uint64_t mw = g.dmask[(static_cast<size_t>(c) * nd + k) * nw + w];
                for (int b = 0; b < 64; ++b)
                    if ((mw >> b) & 1ull) {
                        const int32_t t = static_cast<int32_t>(w * 64) + b;
                        got.insert({t - g.doff[k], t});
                    }
            }
        }
        if (got != want) fail("lists and masks differ from the folded relation", n, c);
        if (!g.fmask.empty() && !want.count({start, start})) fail("first masks without a loop", n, c);
    }
}

void check_nfa(const Nfa &a, Rng &r, int inputs, int max_len, int alphabet,
               DiagMode mode = DiagMode::automatic) {
    validate_nfa(a.in());
    GeneralHost g = build_general(a.in(), mode);
    if (g.ok && g.gather_ok && a.n <= 5000) check_gather_partition(g, a.start, a.n);
    if (mode == DiagMode::off && (g.ndiag || !g.fmask.empty())) fail("masks built when off", a.n, 0);
    if (!g.ok) fail("fold refused a small automaton", a.n, 0);
    if (static_cast<int32_t>(g.rows.row.size()) != a.n + 1) fail("folded rows size", a.n, 0);
    if (g.gather_ok != (a.n <= 65536)) fail("gather lists presence", a.n, 0);
    // the folded rows hold no duplicate (state, label, target) and no epsilon edges
    for (int32_t s = 0; s < a.n; ++s) {
        std::set<uint32_t> seen;
        for (int32_t e = g.rows.row[s]; e < g.rows.row[s + 1]; ++e)
            if (!seen.insert(g.rows.edge[e]).second) fail("duplicate folded edge", s, e);
    }
    if (g.rows.has_eps) fail("folded rows keep epsilon edges", a.n, 0);
    std::vector<uint8_t> data;
    std::vector<int64_t> off{0};
    for (int k = 0; k < inputs; ++k) {
        const int len = r.below(max_len + 1);
        for (int q = 0; q < len; ++q) data.push_back(static_cast<uint8_t>(r.below(alphabet + 1)));
        off.push_back(static_cast<int64_t>(data.size()));
    }
    data.push_back(0);
    std::vector<int32_t> f(inputs), m(inputs);
    std::vector<int64_t> ex(inputs);
    if (oracle_nfa_batch(a.n, a.start, a.acc.data(), a.srp.data(), a.st.data(), a.ss.data(),
                         a.erp.data(), a.et.data(), data.data(), off.data(), inputs, f.data(),
                         m.data(), ex.data(), 1, 2) != 0)
        fail("oracle_nfa_batch failed");
    for (int k = 0; k < inputs; ++k) {
        const uint8_t *p = data.data() + off[k];
        const int len = static_cast<int>(off[k + 1] - off[k]);
        const auto want = sim_ref(a, p, len);
        if (want != std::make_pair(f[k], m[k])) fail("oracle vs set simulation", a.n, k);
        if (sim_rows(g.rows, a.acc.data(), p, len) != want) fail("folded rows", a.n, k);
        if (g.gather_ok && sim_gather(g, p, len) != want) fail("gather lists", a.n, k);
        ++g_cases;
    }
}

void check_fold_cap() {
    // a 3000-state epsilon cycle: every closure is every state (9 M entries)
    Nfa a;
    a.n = 3000;
    a.acc.assign(a.n, 0);
    a.srp.assign(a.n + 1, 0);
    a.erp.push_back(0);
    for (int32_t s = 0; s < a.n; ++s) {
        a.et.push_back((s + 1) % a.n);
        a.erp.push_back(s + 1);
    }
    a.st.resize(1);
    a.ss.resize(1);
    FoldedHost f = fold_epsilon(a.in());
    if (f.ok || f.why.empty()) fail("fold cap not enforced");
    GeneralHost g = build_general(a.in());
    if (g.ok) fail("general layout built past the fold cap");
    // a small cap refuses a modest blow-up, a large one accepts it
    Rng r(77);
    Nfa b = random_nfa(r, 200, 8, 3, 3, 0.05, 0.1);
    if (fold_epsilon(b.in(), 10).ok) fail("cap of 10 edges accepted");
    if (!fold_epsilon(b.in(), 1 << 20).ok) fail("cap of 1M edges refused");
    ++g_cases;
}

void check_dfa(Rng &r, int32_t n, double p_acc) {
    std::vector<int32_t> trans(static_cast<size_t>(n) * 256);
    for (auto &t : trans) t = r.below(n);
    std::vector<uint8_t> acc(n);
    for (auto &x : acc) x = r.unit() < p_acc;
    const int32_t start = r.below(n);
    DfaHost h = build_dfa(n, start, acc.data(), trans.data());
    if (h.esz != (n <= 256 ? 1u : (n <= 0x7fff ? 2u : 4u)) || h.wide != (h.esz == 4))
        fail("DFA entry width", n, 0);
    const int inputs = 64;
    std::vector<uint8_t> data;
    std::vector<int64_t> off{0};
    for (int k = 0; k < inputs; ++k) {
        const int len = r.below(600);
        for (int q = 0; q < len; ++q) data.push_back(static_cast<uint8_t>(r.below(256)));
        off.push_back(static_cast<int64_t>(data.size()));
    }
    data.push_back(0);
    std::vector<int32_t> f(inputs), m(inputs);
    std::vector<int64_t> ex(inputs);
    if (oracle_dfa_batch(n, start, acc.data(), trans.data(), data.data(), off.data(), inputs,
                         f.data(), m.data(), ex.data(), 2) != 0)
        fail("oracle_dfa_batch failed");
    const uint32_t esz = h.esz;
    const size_t colen = esz == 1 ? 256 : static_cast<size_t>(n);
    for (int k = 0; k < inputs; ++k) {
        // the kernels' walk: e = table[byte * col + e], accepting iff e >= accept_from
        uint32_t e = h.start_entry;
        std::pair<int, int> got{0, 0};
        if (h.start_accepts) {
            got = {1, 0};
        } else {
            for (int64_t q = off[k]; q < off[k + 1]; ++q) {
                const size_t at = static_cast<size_t>(data[q]) * colen + e / esz;
                e = esz == 1 ? h.t8[at] : (esz == 2 ? h.t16[at] : h.t32[at]);
                if (e >= h.accept_from) {
                    got = {1, static_cast<int>(q - off[k] + 1)};
                    break;
                }
            }
        }
        if (got != std::make_pair(f[k], m[k])) fail("DFA table walk", n, k);
        ++g_cases;
    }
}

void check_tiles(Rng &r) {
    // a union of chains behind a start with an ANY self-loop
    Nfa a;
    const int chains = 40;
    a.n = 1;
    std::vector<std::pair<int32_t, int32_t>> first;  // (state, label) of first positions
    std::vector<std::vector<std::pair<int32_t, int32_t>>> out(1);
    out[0].push_back({0, kAnySymbol});
    for (int c = 0; c < chains; ++c) {
        const int len = 1 + r.below(12);
        const int32_t head = a.n;
        for (int q = 0; q < len; ++q) out.emplace_back();
        a.n += len;
        const int lab0 = r.below(256);
        out[0].push_back({head, lab0});
        for (int q = 0; q + 1 < len; ++q) out[head + q].push_back({head + q + 1, r.below(256)});
    }
    a.start = 0;
    a.acc.assign(a.n, 0);
    a.srp.push_back(0);
    for (int32_t s = 0; s < a.n; ++s) {
        for (auto [t, l] : out[s]) {
            a.st.push_back(t);
            a.ss.push_back(l);
        }
        a.srp.push_back(static_cast<int32_t>(a.st.size()));
    }
    a.erp.assign(a.n + 1, 0);
    a.et.resize(1);
    TileHost t = build_tiles(a.in());
    if (!t.ok) fail("pattern union refused by the tiles");
    std::set<int32_t> slots;
    for (int32_t s = 1; s < a.n; ++s) {
        const int32_t sl = t.slot_of_state[s];
        if (sl < 0 || !slots.insert(sl).second) fail("tile slot not a bijection", s, sl);
    }
    if (!t.start_always) fail("ANY self-loop not seen");
    // an epsilon edge, then a state with two successors: both refused with a reason
    Nfa e = a;
    e.erp.assign(a.n + 1, 0);
    for (int32_t s = 1; s <= a.n; ++s) e.erp[s] = 1;
    e.et.assign(1, 1);
    if (build_tiles(e.in()).ok) fail("tiles accepted an epsilon edge");
    Nfa two = a;
    for (int32_t s = 2; s <= a.n; ++s) two.srp[s] += 2;  // state 1 gets two extra edges
    {
        std::vector<int32_t> st2, ss2;
        for (int32_t s = 0; s < a.n; ++s) {
            for (int32_t k = a.srp[s]; k < a.srp[s + 1]; ++k) {
                st2.push_back(a.st[k]);
                ss2.push_back(a.ss[k]);
            }
            if (s == 1) {
                st2.push_back(2);
                ss2.push_back(5);
                st2.push_back(3);
                ss2.push_back(6);
            }
        }
        two.st = st2;
        two.ss = ss2;
    }
    TileHost bad = build_tiles(two.in());
    if (bad.ok || bad.why.empty()) fail("tiles accepted a state with two successors");
    ++g_cases;
}

// The byte-triple lists of the chain lane tier: every header against a brute-force
// enumeration from the automaton's own edges (first position u on c1, its successor v on c2,
// v's successor w on c3 gives slot(w)), the pair accept bits likewise, and a walk of random
// inputs the way the lane tier walks them (held slots continued through `next`, depth-3
// arrivals from the lists, depth-1 and depth-2 accepts from the flags) against the set
// simulation.
void check_triples(Rng &r) {
    Nfa a;
    a.n = 1;
    std::vector<std::vector<std::pair<int32_t, int32_t>>> out(1);
    out[0].push_back({0, kAnySymbol});
    struct Pat {
        int32_t head;
        std::vector<std::vector<int>> cls;
    };
    std::vector<Pat> pats;
    auto add_pattern = [&](std::vector<std::vector<int>> cls, bool accepting) {
        const int32_t head = a.n;
        const int len = static_cast<int>(cls.size());
        for (int q = 0; q < len; ++q) out.emplace_back();
        a.n += len;
        for (int c : cls[0]) out[0].push_back({head, c});
        for (int q = 0; q + 1 < len; ++q)
            for (int c : cls[q + 1]) out[head + q].push_back({head + q + 1, c});
        a.acc.resize(a.n, 0);
        if (accepting) a.acc[head + len - 1] = 1;
        pats.push_back({head, std::move(cls)});
    };
    a.acc.assign(1, 0);
    for (int p = 0; p < 80; ++p) {
        const int len = 1 + r.below(8);
        std::vector<std::vector<int>> cls(len);
        for (auto &c : cls) {
            std::set<int> bytes;
            const int w = 1 + r.below(5);
            while (static_cast<int>(bytes.size()) < w) bytes.insert(r.below(16));
            c.assign(bytes.begin(), bytes.end());
        }
        add_pattern(cls, r.below(4) != 0);
    }
    for (int p = 0; p < 9; ++p)  // one list of 9 entries: (1, 2, 3)
        add_pattern({{1}, {2}, {3}, {4 + p % 3}, {5}}, p % 2 == 0);
    a.start = 0;
    a.srp.assign(1, 0);
    for (int32_t s = 0; s < a.n; ++s) {
        for (auto [t, l] : out[s]) {
            a.st.push_back(t);
            a.ss.push_back(l);
        }
        a.srp.push_back(static_cast<int32_t>(a.st.size()));
    }
    a.erp.assign(a.n + 1, 0);
    a.et.resize(1);
    const TileHost t = build_tiles(a.in());
    if (!t.ok || !t.start_always || t.pair.empty()) fail("triple test automaton refused");
    const TripleHost tr = build_triples(t);
    if (!tr.ok) fail("byte-triple lists refused");
    // brute force over the patterns
    std::vector<std::vector<int32_t>> want(size_t(1) << 24);
    std::vector<uint8_t> want_acc(size_t(1) << 24, 0), want_pacc(65536, 0);
    int64_t entries = 0;
    for (const Pat &pt : pats) {
        const int len = static_cast<int>(pt.cls.size());
        if (len >= 2)
            for (int c1 : pt.cls[0])
                for (int c2 : pt.cls[1])
                    if (a.acc[pt.head + 1]) want_pacc[c1 * 256 + c2] = 1;
        if (len < 3) continue;
        const int32_t w = t.slot_of_state[pt.head + 2];
        for (int c1 : pt.cls[0])
            for (int c2 : pt.cls[1])
                for (int c3 : pt.cls[2]) {
                    const size_t i = (static_cast<size_t>(c1) << 16) | (c2 << 8) | c3;
                    want[i].push_back(w);
                    if (a.acc[pt.head + 2]) want_acc[i] = 1;
                    ++entries;
                }
    }
    if (entries != tr.entries) fail("triple entry count", static_cast<int>(entries),
                                    static_cast<int>(tr.entries));
    for (size_t i = 0; i < want.size(); ++i) {
        const uint32_t h = tr.hdr[i];
        const uint32_t cnt = (h >> 24) & 127u;
        if (cnt != want[i].size() || (h >> 31) != want_acc[i])
            fail("triple header", static_cast<int>(i), static_cast<int>(cnt));
        if (((tr.nonempty[i >> 5] >> (i & 31)) & 1u) != (cnt > 0 ? 1u : 0u))
            fail("triple non-empty bit", static_cast<int>(i));
        if (cnt == 0) continue;
        std::vector<int32_t> got;
        if (cnt == 1) {
            got.push_back(static_cast<int32_t>(h & 0xffffffu));
        } else {
            const uint32_t o = h & 0xffffffu;
            if (o & 3u) fail("triple list not 4-aligned", static_cast<int>(i));
            for (uint32_t k = 0; k < cnt; ++k) got.push_back(tr.slot.at(o + k));
        }
        std::vector<int32_t> w = want[i];
        std::sort(w.begin(), w.end());
        std::sort(got.begin(), got.end());
        if (got != w) fail("triple list", static_cast<int>(i), static_cast<int>(cnt));
    }
    for (int i = 0; i < 65536; ++i)
        if (((tr.pair_acc[i >> 5] >> (i & 31)) & 1u) != want_pacc[i]) fail("pair accept bit", i);
    // the lane tier's walk
    const size_t words = static_cast<size_t>(t.tiles + 1) * kTileWords;
    auto nextbit = [&](int c, uint32_t s) {
        return (t.next[static_cast<size_t>(c) * words + (s >> 6)] >> (s & 63)) & 1ull;
    };
    auto accbit = [&](uint32_t s) { return (t.acc[s >> 6] >> (s & 63)) & 1ull; };
    for (int k = 0; k < 400; ++k) {
        const int len = r.below(200);
        std::vector<uint8_t> in(static_cast<size_t>(len));
        for (auto &b : in) b = static_cast<uint8_t>(r.below(17));
        std::vector<uint32_t> held, nxt;
        std::pair<int, int> got{0, 0};
        for (int j = 0; j < len && !got.first; ++j) {
            const int b = in[static_cast<size_t>(j)];
            bool hit = false;
            nxt.clear();
            for (uint32_t h : held)
                if (nextbit(b, h + 1)) {
                    hit |= accbit(h + 1) != 0;
                    nxt.push_back(h + 1);
                }
            if (j >= 2) {
                const uint32_t h = tr.hdr[(static_cast<size_t>(in[j - 2]) << 16) | (in[j - 1] << 8) | b];
                const uint32_t cnt = (h >> 24) & 127u;
                hit |= (h >> 31) != 0;
                for (uint32_t m = 0; m < cnt; ++m)
                    nxt.push_back(cnt == 1 ? (h & 0xffffu) : tr.slot[(h & 0xffffffu) + m]);
            }
            hit |= (t.bflags[b] & 2u) != 0;
            if (j >= 1) hit |= ((tr.pair_acc[(in[j - 1] << 3) | (b >> 5)] >> (b & 31)) & 1u) != 0;
            if (hit) got = {1, j + 1};
            held.swap(nxt);
        }
        if (got != sim_ref(a, in.data(), len)) fail("lane-tier walk", k, got.second);
        ++g_cases;
    }
    // refusals: no ANY self-loop on the start state
    Nfa no_loop = a;
    no_loop.ss[0] = 0;
    const TileHost t2 = build_tiles(no_loop.in());
    if (t2.ok && build_triples(t2).ok) fail("byte-triple lists without an always-active start");
    ++g_cases;
}

void check_validation() {
    Nfa a;
    a.n = 3;
    a.start = 0;
    a.acc.assign(3, 0);
    a.srp = {0, 1, 1, 1};
    a.st = {5};  // target out of range
    a.ss = {7};
    a.erp = {0, 0, 0, 0};
    a.et.resize(1);
    bool threw = false;
    try {
        validate_nfa(a.in());
    } catch (const Error &e) {
        threw = e.code == 1;
    }
    if (!threw) fail("out-of-range target accepted");
    a.st = {1};
    a.ss = {257};  // label past ANY
    threw = false;
    try {
        validate_nfa(a.in());
    } catch (const Error &e) {
        threw = e.code == 1;
    }
    if (!threw) fail("label 257 accepted");
    a.ss = {256};
    a.srp = {0, 1, 0, 1};  // decreasing row pointer
    threw = false;
    try {
        validate_nfa(a.in());
    } catch (const Error &e) {
        threw = e.code == 1;
    }
    if (!threw) fail("decreasing row pointers accepted");
    std::vector<int32_t> trans(256 * 2, 0);
    trans[17] = 2;  // DFA target out of range
    std::vector<uint8_t> acc(2, 0);
    threw = false;
    try {
        build_dfa(2, 0, acc.data(), trans.data());
    } catch (const Error &e) {
        threw = e.code == 1;
    }
    if (!threw) fail("DFA target out of range accepted");
    ++g_cases;
}

}  // namespace

int main() {
    Rng r(2026);
    // small automata: every technique's layout, alphabet 6 so that runs survive
    for (int trial = 0; trial < 300; ++trial) {
        const int32_t n = 1 + r.below(trial < 200 ? 64 : 600);
        Nfa a = random_nfa(r, n, 6, 3, 2, 0.05, 0.1);
        check_nfa(a, r, 24, 40, 6);
    }
    // multi-word sets and big rows (the per-label index), ANY-heavy rows
    for (int32_t n : {65, 300, 2000, 20000}) {
        Nfa a = random_nfa(r, n, 4, 24, 1, 0.02, 0.01);
        check_nfa(a, r, 16, 60, 4);
    }
    {
        Nfa a = random_nfa(r, 70000, 8, 3, 1, 0.01, 0.001);  // above 65536: no gather lists
        check_nfa(a, r, 8, 30, 8);
    }
    // first and diagonal masks: pattern unions with optional positions (chains forward and
    // reversed, start looping on ANY or on each byte), masks chosen by the layout, forced
    // and switched off; random automata with forced masks (any offsets, both signs)
    for (int trial = 0; trial < 24; ++trial) {
        const bool any_loop = trial % 2 == 0, reversed = (trial / 2) % 2 == 1;
        const int patterns = trial < 12 ? 3 + r.below(20) : 150 + r.below(200);
        Nfa a = union_nfa(r, patterns, 2, 9, 4, 8, 0.3, any_loop, reversed);
        const DiagMode mode = trial % 3 == 0 ? DiagMode::automatic
                                             : (trial % 3 == 1 ? DiagMode::force : DiagMode::off);
        check_nfa(a, r, 20, 80, 8, mode);
    }
    {
        // about 200 edges per byte on offset 1 and 48 from the start, against 41 words
        Nfa a = union_nfa(r, 400, 4, 9, 60, 256, 0.25, true, false);
        GeneralHost g = build_general(a.in(), DiagMode::automatic);
        if (g.fmask.empty() || g.ndiag == 0 || g.doff[0] != 1) fail("union: masks not chosen", a.n, g.ndiag);
    }
    for (int trial = 0; trial < 60; ++trial) {
        const int32_t n = 1 + r.below(trial < 40 ? 130 : 700);
        Nfa a = random_nfa(r, n, 6, 3, 2, 0.05, 0.1);
        check_nfa(a, r, 16, 40, 6, DiagMode::force);
    }
    check_fold_cap();
    for (int32_t n : {1, 2, 255, 256, 300, 4096, 33000}) check_dfa(r, n, n < 300 ? 0.1 : 0.01);
    check_tiles(r);
    check_triples(r);
    check_validation();
    std::printf("host_check ok: %d cases\n", g_cases);
    return 0;
}
