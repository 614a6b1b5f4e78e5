// Host-side preprocessing of upstream CSR automata into the device layouts (see layout.hpp).
#include "layout.hpp"

#include <algorithm>
#include <climits>
#include <numeric>
#include <string>

#include "internal.hpp"

namespace gfsm {

namespace {

[[noreturn]] void bad(const std::string &m) { throw Error(1, m); }

struct Bits256 {
    uint64_t w[4] = {0, 0, 0, 0};
    void add_label(int32_t lab) {
        if (lab == kAnySymbol) {
            w[0] = w[1] = w[2] = w[3] = ~0ull;
        } else {
            w[lab >> 6] |= 1ull << (lab & 63);
        }
    }
    bool any() const { return (w[0] | w[1] | w[2] | w[3]) != 0; }
    bool operator==(const Bits256 &o) const {
        return w[0] == o.w[0] && w[1] == o.w[1] && w[2] == o.w[2] && w[3] == o.w[3];
    }
    bool has(int b) const { return (w[b >> 6] >> (b & 63)) & 1ull; }
    int count() const {
        return __builtin_popcountll(w[0]) + __builtin_popcountll(w[1]) +
               __builtin_popcountll(w[2]) + __builtin_popcountll(w[3]);
    }
    std::vector<int> bytes() const {
        std::vector<int> out;
        for (int c = 0; c < 256; ++c)
            if (has(c)) out.push_back(c);
        return out;
    }
};

}  // namespace

void validate_nfa(const NfaInput &a) {
    if (a.n < 1) bad("num_states must be >= 1");
    if (a.n >= kMaxPackedStates)
        bad("num_states " + std::to_string(a.n) + " exceeds the supported maximum " +
            std::to_string(kMaxPackedStates - 1));
    if (a.start < 0 || a.start >= a.n) bad("start_state out of range");
    if (!a.accept || !a.srp || !a.erp) bad("null array");
    if (a.srp[0] != 0 || a.erp[0] != 0) bad("row pointers must start at 0");
    for (int32_t s = 0; s < a.n; ++s) {
        if (a.srp[s + 1] < a.srp[s]) bad("sym_row_ptr is not non-decreasing");
        if (a.erp[s + 1] < a.erp[s]) bad("eps_row_ptr is not non-decreasing");
    }
    const int32_t nnz = a.srp[a.n], nne = a.erp[a.n];
    if (nnz > 0 && (!a.st || !a.ss)) bad("null symbol-edge array");
    if (nne > 0 && !a.et) bad("null epsilon-edge array");
    for (int32_t k = 0; k < nnz; ++k) {
        if (a.st[k] < 0 || a.st[k] >= a.n) bad("sym_targets entry out of range");
        if (a.ss[k] < 0 || a.ss[k] > kAnySymbol) bad("sym_symbols entry outside 0..256");
    }
    for (int32_t k = 0; k < nne; ++k)
        if (a.et[k] < 0 || a.et[k] >= a.n) bad("eps_targets entry out of range");
}

WorklistHost build_worklist(const NfaInput &a) {
    WorklistHost h;
    const int32_t n = a.n;
    const int32_t nnz = a.srp[n];
    h.nwords = (n + 63) / 64;
    h.row.assign(a.srp, a.srp + n + 1);
    h.edge.resize(static_cast<size_t>(nnz));
    h.bigidx.assign(static_cast<size_t>(n), -1);
    int32_t count[kAnySymbol + 2];
    for (int32_t s = 0; s < n; ++s) {
        const int32_t lo = a.srp[s], hi = a.srp[s + 1];
        std::fill(count, count + kAnySymbol + 2, 0);
        for (int32_t k = lo; k < hi; ++k) count[a.ss[k] + 1]++;
        for (int c = 0; c <= kAnySymbol; ++c) count[c + 1] += count[c];
        if (hi - lo > kBigRow) {
            h.bigidx[s] = static_cast<int32_t>(h.bigoff.size() / kBigStride);
            for (int c = 0; c <= kAnySymbol; ++c) h.bigoff.push_back(lo + count[c]);
            h.bigoff.push_back(hi);
        }
        for (int32_t k = lo; k < hi; ++k) {
            const int32_t pos = lo + count[a.ss[k]]++;
            h.edge[pos] = (static_cast<uint32_t>(a.st[k]) << kLabelBits) |
                          static_cast<uint32_t>(a.ss[k]);
        }
    }
    h.erow.assign(a.erp, a.erp + n + 1);
    h.etgt.assign(a.et, a.et + a.erp[n]);
    h.has_eps = a.erp[n] > 0;
    h.accw.assign(static_cast<size_t>(h.nwords), 0ull);
    for (int32_t s = 0; s < n; ++s)
        if (a.accept[s]) h.accw[s >> 6] |= 1ull << (s & 63);
    // start closure: depth-first over epsilon edges, seeds included
    std::vector<uint8_t> seen(static_cast<size_t>(n), 0);
    std::vector<int32_t> stack{a.start};
    seen[a.start] = 1;
    h.start_list.push_back(a.start);
    while (!stack.empty()) {
        const int32_t s = stack.back();
        stack.pop_back();
        for (int32_t k = a.erp[s]; k < a.erp[s + 1]; ++k) {
            const int32_t t = a.et[k];
            if (!seen[t]) {
                seen[t] = 1;
                h.start_list.push_back(t);
                stack.push_back(t);
            }
        }
    }
    for (int32_t s : h.start_list)
        if (a.accept[s]) h.start_accepts = true;
    return h;
}

TileHost build_tiles(const NfaInput &a) {
    TileHost h;
    const int32_t n = a.n, s0 = a.start;
    if (a.erp[n] > 0) {
        h.why = "epsilon edges";
        return h;
    }
    std::vector<int32_t> pred(static_cast<size_t>(n), -1), succ(static_cast<size_t>(n), -1);
    std::vector<Bits256> from_start(static_cast<size_t>(n)), from_pred(static_cast<size_t>(n));
    Bits256 self;
    for (int32_t s = 0; s < n; ++s) {
        for (int32_t k = a.srp[s]; k < a.srp[s + 1]; ++k) {
            const int32_t t = a.st[k], lab = a.ss[k];
            if (t == s0) {
                if (s != s0) {
                    h.why = "edge into the start state";
                    return h;
                }
                self.add_label(lab);
                continue;
            }
            if (s == s0) {
                from_start[t].add_label(lab);
                continue;
            }
            if (succ[s] != -1 && succ[s] != t) {
                h.why = "a state with two successors";
                return h;
            }
            succ[s] = t;
            if (pred[t] != -1 && pred[t] != s) {
                h.why = "a state with two predecessors";
                return h;
            }
            pred[t] = s;
            from_pred[t].add_label(lab);
        }
    }
    for (int32_t t = 0; t < n; ++t) {
        if (t == s0) continue;
        if (pred[t] != -1 && from_start[t].any() && !(from_start[t] == from_pred[t])) {
            h.why = "inhomogeneous labels";
            return h;
        }
    }
    // chains: heads have no non-start predecessor; every non-start state must be on one
    std::vector<std::vector<int32_t>> chains;
    std::vector<uint8_t> placed(static_cast<size_t>(n), 0);
    size_t covered = 0;
    for (int32_t t = 0; t < n; ++t) {
        if (t == s0 || pred[t] != -1) continue;
        std::vector<int32_t> c;
        for (int32_t u = t; u != -1; u = succ[u]) {
            if (placed[u]) {
                h.why = "a cycle";
                return h;
            }
            placed[u] = 1;
            c.push_back(u);
            if (static_cast<int32_t>(c.size()) > kTileStates) {
                h.why = "a chain longer than one tile";
                return h;
            }
        }
        covered += c.size();
        chains.push_back(std::move(c));
    }
    if (covered != static_cast<size_t>(n - 1)) {
        h.why = "a cycle";
        return h;
    }
    // best-fit decreasing into 4096-state tiles
    std::vector<size_t> order(chains.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return chains[x].size() > chains[y].size(); });
    std::vector<int32_t> fill;
    std::vector<std::vector<size_t>> bins;
    for (size_t ci : order) {
        const int32_t len = static_cast<int32_t>(chains[ci].size());
        h.max_chain = std::max(h.max_chain, len);
        // best fit: the fullest tile that still has room
        size_t b = fill.size();
        for (size_t t = 0; t < fill.size(); ++t)
            if (fill[t] + len <= kTileStates && (b == fill.size() || fill[t] > fill[b])) b = t;
        if (b == fill.size()) {
            fill.push_back(0);
            bins.emplace_back();
        }
        bins[b].push_back(ci);
        fill[b] += len;
    }
    h.tiles = std::max<int32_t>(1, static_cast<int32_t>(bins.size()));
    const size_t T = static_cast<size_t>(h.tiles);
    h.mask.assign(T * 256 * kTileWords, 0ull);
    h.first.assign(T * kTileWords, 0ull);
    h.hassucc.assign(T * kTileWords, 0ull);
    h.acc.assign((T + 1) * kTileWords, 0ull);  // one all-zero tile past the last
    h.slot_of_state.assign(static_cast<size_t>(n), -1);
    for (size_t b = 0; b < bins.size(); ++b) {
        int32_t slot = 0;
        for (size_t ci : bins[b]) {
            for (int32_t u : chains[ci]) {
                const size_t w = b * kTileWords + (slot >> 6);
                const uint64_t bit = 1ull << (slot & 63);
                h.slot_of_state[u] = static_cast<int32_t>(b) * kTileStates + slot;
                const Bits256 &cls = pred[u] != -1 ? from_pred[u] : from_start[u];
                for (int c = 0; c < 256; ++c)
                    if (cls.has(c)) h.mask[(b * 256 + c) * kTileWords + (slot >> 6)] |= bit;
                if (from_start[u].any()) h.first[w] |= bit;
                if (succ[u] != -1) h.hassucc[w] |= bit;
                if (a.accept[u]) h.acc[w] |= bit;
                ++slot;
            }
        }
    }
    // per-byte lists of the start state's successors (in slot order)
    std::vector<std::vector<int32_t>> by_byte(256);
    h.first_accepts.assign(256, 0u);
    for (int32_t u = 0; u < n; ++u) {
        if (u == s0 || !from_start[u].any()) continue;
        for (int c = 0; c < 256; ++c) {
            if (!from_start[u].has(c)) continue;
            by_byte[c].push_back(h.slot_of_state[u]);
            if (a.accept[u]) h.first_accepts[c] = 1u;
        }
    }
    h.first_off.assign(257, 0);
    for (int c = 0; c < 256; ++c) {
        std::sort(by_byte[c].begin(), by_byte[c].end());
        h.first_off[c + 1] = h.first_off[c] + static_cast<int32_t>(by_byte[c].size());
    }
    h.first_slot.reserve(static_cast<size_t>(h.first_off[256]));
    for (int c = 0; c < 256; ++c)
        h.first_slot.insert(h.first_slot.end(), by_byte[c].begin(), by_byte[c].end());
    for (int c = 0; c < 256; ++c) {
        if (h.first_off[c + 1] > h.first_off[c]) h.first_any[c >> 6] |= 1ull << (c & 63);
        if (h.first_accepts[c]) h.first_acc[c >> 6] |= 1ull << (c & 63);
    }
    // byte-pair lists of second positions (counting pass, then fill)
    {
        std::vector<int32_t> firsts;  // first positions with a chain successor
        int64_t total = 0;
        for (int32_t u = 0; u < n; ++u) {
            if (u == s0 || !from_start[u].any() || succ[u] == -1) continue;
            firsts.push_back(u);
            total += static_cast<int64_t>(from_start[u].count()) * from_pred[succ[u]].count();
        }
        if (total <= kMaxPairSlots) {
            std::vector<int64_t> cnt(65536, 0);
            for (int32_t u : firsts) {
                const std::vector<int> c2s = from_pred[succ[u]].bytes();
                for (int c : from_start[u].bytes())
                    for (int c2 : c2s) ++cnt[c * 256 + c2];
            }
            h.pair.assign(257 * 256 * 2, 0);
            std::vector<int64_t> fill(65536, 0);
            int64_t o = 0;
            for (int i = 0; i < 65536; ++i) {
                fill[i] = o;
                h.pair[2 * i] = static_cast<int32_t>(o);
                h.pair[2 * i + 1] = static_cast<int32_t>(cnt[i]);
                o += cnt[i];
            }
            // 64 padding entries: a wave may read up to 63 past the end of a list
            h.pair_slot.assign(static_cast<size_t>(total) + 64, 0);
            for (int32_t u : firsts) {
                const int32_t v = succ[u];
                const int32_t sv = h.slot_of_state[v];
                const std::vector<int> c2s = from_pred[v].bytes();
                for (int c : from_start[u].bytes())
                    for (int c2 : c2s) {
                        const int i = c * 256 + c2;
                        h.pair_slot[static_cast<size_t>(fill[i]++)] = sv;
                        if (a.accept[v]) h.pair[2 * i + 1] |= INT32_MIN;
                    }
            }
            for (int i = 0; i < 65536; ++i)
                std::sort(h.pair_slot.begin() + h.pair[2 * i],
                          h.pair_slot.begin() + h.pair[2 * i] + (h.pair[2 * i + 1] & INT32_MAX));
        }
    }
    h.bflags.assign(256, 0u);
    for (int c = 0; c < 256; ++c)
        h.bflags[c] = ((h.first_any[c >> 6] >> (c & 63)) & 1ull) |
                      (((h.first_acc[c >> 6] >> (c & 63)) & 1ull) << 1) |
                      (static_cast<uint32_t>(self.has(c)) << 2);
    h.next.assign(256 * (T + 1) * kTileWords, 0ull);
    for (size_t b = 0; b < T; ++b) {
        const uint64_t *hs = &h.hassucc[b * kTileWords];
        for (int c = 0; c < 256; ++c) {
            const uint64_t *m = &h.mask[(b * 256 + c) * kTileWords];
            uint64_t *o = &h.next[(static_cast<size_t>(c) * (T + 1) + b) * kTileWords];
            for (int w = 0; w < kTileWords; ++w) {
                const uint64_t shifted = (hs[w] << 1) | (w > 0 ? hs[w - 1] >> 63 : 0ull);
                o[w] = m[w] & shifted;
            }
        }
    }
    for (int i = 0; i < 4; ++i) h.selfmask[i] = self.w[i];
    h.start_always = (self.w[0] & self.w[1] & self.w[2] & self.w[3]) == ~0ull;
    h.ok = true;
    return h;
}

TripleHost build_triples(const TileHost &t) {
    TripleHost h;
    if (!t.ok) {
        h.why = "not chain-structured";
        return h;
    }
    if (!t.start_always) {
        h.why = "the start state does not loop on every byte";
        return h;
    }
    if (t.pair.empty()) {
        h.why = "no byte-pair lists";
        return h;
    }
    if (static_cast<int64_t>(t.tiles) * kTileStates > 65536) {
        h.why = "more than 65536 slots";
        return h;
    }
    const size_t T = static_cast<size_t>(t.tiles);
    const size_t words = (T + 1) * kTileWords;  // one byte's plane of `next`, in u64 words
    const size_t S = words * 64;                 // slots, the all-zero tile included
    // bytes on which slot s is entered from slot s - 1 (bit s of next[c]), as a CSR
    std::vector<int32_t> boff(S + 1, 0);
    for (int c = 0; c < 256; ++c) {
        const uint64_t *plane = &t.next[static_cast<size_t>(c) * words];
        for (size_t w = 0; w < words; ++w)
            for (uint64_t m = plane[w]; m; m &= m - 1)
                ++boff[w * 64 + static_cast<size_t>(__builtin_ctzll(m)) + 1];
    }
    for (size_t s = 0; s < S; ++s) boff[s + 1] += boff[s];
    std::vector<uint8_t> bytes(static_cast<size_t>(boff[S]));
    {
        std::vector<int32_t> fill(boff.begin(), boff.end() - 1);
        for (int c = 0; c < 256; ++c) {
            const uint64_t *plane = &t.next[static_cast<size_t>(c) * words];
            for (size_t w = 0; w < words; ++w)
                for (uint64_t m = plane[w]; m; m &= m - 1)
                    bytes[static_cast<size_t>(fill[w * 64 + static_cast<size_t>(__builtin_ctzll(m))]++)] =
                        static_cast<uint8_t>(c);
        }
    }
    auto accepting = [&](uint32_t s) { return (t.acc[s >> 6] >> (s & 63)) & 1ull; };
    // total entries first, so that an automaton over the cap costs no large allocation
    int64_t total = 0;
    for (int i = 0; i < 65536; ++i) {
        const int32_t o = t.pair[2 * i], cnt = t.pair[2 * i + 1] & INT32_MAX;
        for (int32_t e = o; e < o + cnt; ++e) {
            const size_t w = static_cast<size_t>(t.pair_slot[static_cast<size_t>(e)]) + 1;
            total += boff[w + 1] - boff[w];
        }
    }
    if (total > kMaxTripleSlots) {
        h.why = "byte-triple lists of " + std::to_string(total) + " entries (cap " +
                std::to_string(kMaxTripleSlots) + ")";
        return h;
    }
    h.entries = total;
    h.hdr.assign(size_t(1) << 24, 0u);
    h.pair_acc.assign(2048, 0u);
    // one pair (c1, c2) at a time: its 256 triples are one contiguous block of headers
    uint32_t cnt3[256], acc3[256], pos3[256];
    for (int i = 0; i < 65536; ++i) {
        const int32_t o = t.pair[2 * i], cnt = t.pair[2 * i + 1] & INT32_MAX;
        if (t.pair[2 * i + 1] < 0) h.pair_acc[i >> 5] |= 1u << (i & 31);
        if (cnt == 0) continue;
        std::fill(cnt3, cnt3 + 256, 0u);
        std::fill(acc3, acc3 + 256, 0u);
        for (int32_t e = o; e < o + cnt; ++e) {
            const uint32_t w = static_cast<uint32_t>(t.pair_slot[static_cast<size_t>(e)]) + 1;
            for (int32_t k = boff[w]; k < boff[w + 1]; ++k) {
                ++cnt3[bytes[static_cast<size_t>(k)]];
                if (accepting(w)) acc3[bytes[static_cast<size_t>(k)]] = 1u;
            }
        }
        for (int c3 = 0; c3 < 256; ++c3) {
            if (cnt3[c3] > 127) {
                h = TripleHost();
                h.why = "a byte-triple list of more than 127 entries";
                return h;
            }
            if (cnt3[c3] >= 2) {
                pos3[c3] = static_cast<uint32_t>(h.slot.size());
                h.slot.resize(h.slot.size() + ((cnt3[c3] + 3) & ~3u), 0);
                if (h.slot.size() > static_cast<size_t>(kMaxTripleSlots)) {
                    h = TripleHost();
                    h.why = "padded byte-triple lists above the offset range";
                    return h;
                }
            }
        }
        uint32_t *blk = &h.hdr[static_cast<size_t>(i) << 8];
        for (int c3 = 0; c3 < 256; ++c3)
            if (cnt3[c3] >= 2) blk[c3] = (acc3[c3] << 31) | (cnt3[c3] << 24) | pos3[c3];
        for (int32_t e = o; e < o + cnt; ++e) {
            const uint32_t w = static_cast<uint32_t>(t.pair_slot[static_cast<size_t>(e)]) + 1;
            for (int32_t k = boff[w]; k < boff[w + 1]; ++k) {
                const int c3 = bytes[static_cast<size_t>(k)];
                if (cnt3[c3] == 1) blk[c3] = (acc3[c3] << 31) | (1u << 24) | w;
                else h.slot[pos3[c3]++] = static_cast<uint16_t>(w);
            }
        }
    }
    h.slot.resize(h.slot.size() + 8, 0);  // tail padding: the 8-byte loads stay in bounds
    h.nonempty.assign(size_t(1) << 19, 0u);
    for (size_t i = 0; i < h.hdr.size(); ++i)
        if (h.hdr[i] >> 24 & 127u) h.nonempty[i >> 5] |= 1u << (i & 31);
    h.ok = true;
    return h;
}

FoldedHost fold_epsilon(const NfaInput &a, int64_t cap) {
    FoldedHost h;
    const int32_t n = a.n;
    const int32_t nnz = a.srp[n], nne = a.erp[n];
    if (cap <= 0) cap = std::max<int64_t>(4ll * nnz, int64_t(1) << 22);
    // closure(t) for every state, depth-first over epsilon edges (seeds included); the
    // total of the closure sizes is bounded by the cap too
    std::vector<int64_t> coff(static_cast<size_t>(n) + 1, 0);
    std::vector<int32_t> clos;
    std::vector<int32_t> stamp(static_cast<size_t>(n), -1), stack;
    for (int32_t t = 0; t < n; ++t) {
        coff[t] = static_cast<int64_t>(clos.size());
        clos.push_back(t);
        if (nne > 0) {
            stamp[t] = t;
            stack.assign(1, t);
            while (!stack.empty()) {
                const int32_t s = stack.back();
                stack.pop_back();
                for (int32_t k = a.erp[s]; k < a.erp[s + 1]; ++k) {
                    const int32_t u = a.et[k];
                    if (stamp[u] != t) {
                        stamp[u] = t;
                        clos.push_back(u);
                        stack.push_back(u);
                    }
                }
            }
        }
        if (static_cast<int64_t>(clos.size()) > cap + n) {
            h.why = "epsilon closures exceed " + std::to_string(cap) + " states in total";
            return h;
        }
    }
    coff[n] = static_cast<int64_t>(clos.size());
    int64_t bound = 0;
    for (int32_t k = 0; k < nnz; ++k) bound += coff[a.st[k] + 1] - coff[a.st[k]];
    if (bound > cap) {
        h.why = "folding the epsilon closure gives up to " + std::to_string(bound) +
                " edges (cap " + std::to_string(cap) + ")";
        return h;
    }
    // folded rows: for each (s, label) the union of the closures of its targets, each
    // target once (first occurrence order)
    h.srp.assign(static_cast<size_t>(n) + 1, 0);
    h.st.reserve(static_cast<size_t>(bound));
    h.ss.reserve(static_cast<size_t>(bound));
    // The row's edges are visited in label order, so all edges of one (s, label) pair are
    // consecutive and one stamp per target (key s * 257 + label) drops every duplicate.
    std::vector<int64_t> seen(static_cast<size_t>(n), -1);
    std::vector<int32_t> order;
    for (int32_t s = 0; s < n; ++s) {
        order.resize(static_cast<size_t>(a.srp[s + 1] - a.srp[s]));
        std::iota(order.begin(), order.end(), a.srp[s]);
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t x, int32_t y) { return a.ss[x] < a.ss[y]; });
        for (int32_t k : order) {
            const int32_t lab = a.ss[k];
            const int64_t key = static_cast<int64_t>(s) * (kAnySymbol + 1) + lab;
            const int32_t t = a.st[k];
            for (int64_t q = coff[t]; q < coff[t + 1]; ++q) {
                const int32_t u = clos[q];
                if (seen[u] == key) continue;
                seen[u] = key;
                h.st.push_back(u);
                h.ss.push_back(lab);
            }
        }
        h.srp[s + 1] = static_cast<int32_t>(h.st.size());
    }
    for (int64_t q = coff[a.start]; q < coff[a.start + 1]; ++q) {
        h.start_set.push_back(clos[q]);
        if (a.accept[clos[q]]) h.start_accepts = true;
    }
    h.ok = true;
    return h;
}

namespace {

// Whether the masks of g carry the folded edge s -> t (whatever its label).
bool mask_covers(const GeneralHost &g, int32_t start, int32_t s, int32_t t) {
    if (!g.fmask.empty() && s == start) return true;
    for (int32_t k = 0; k < g.ndiag; ++k)
        if (t - s == g.doff[k]) return true;
    return false;
}

// First masks and diagonal masks of the gather layout (see GeneralHost).
void build_gather_masks(GeneralHost &g, const FoldedHost &f, int32_t start, int32_t n,
                        DiagMode mode) {
    if (mode == DiagMode::off) return;
    const int32_t nwords = (n + 63) / 64;
    const size_t nw = static_cast<size_t>(nwords);
    // list edges a mask would replace, summed over the 256 bytes (an ANY edge is read on
    // every byte); a mask costs nwords words per byte
    const double keep = mode == DiagMode::force ? 0.5 : 256.0 * nwords;
    auto weight = [](int32_t lab) { return lab == kAnySymbol ? 256.0 : 1.0; };
    // the start stays active when it loops on every byte
    bool loop_any = false;
    std::vector<uint8_t> loops(kAlphabet, 0);
    double start_w = 0;
    for (int32_t k = f.srp[start]; k < f.srp[start + 1]; ++k) {
        start_w += weight(f.ss[k]);
        if (f.st[k] != start) continue;
        if (f.ss[k] == kAnySymbol) loop_any = true;
        else loops[static_cast<size_t>(f.ss[k])] = 1;
    }
    const bool always = loop_any || std::all_of(loops.begin(), loops.end(),
                                                [](uint8_t x) { return x != 0; });
    if (always && start_w >= keep) {
        g.fmask.assign(static_cast<size_t>(kAlphabet) * nw, 0ull);
        for (int32_t k = f.srp[start]; k < f.srp[start + 1]; ++k) {
            const int32_t t = f.st[k];
            const uint64_t bit = 1ull << (t & 63);
            const int c0 = f.ss[k] == kAnySymbol ? 0 : f.ss[k];
            const int c1 = f.ss[k] == kAnySymbol ? kAlphabet : f.ss[k] + 1;
            for (int c = c0; c < c1; ++c) g.fmask[static_cast<size_t>(c) * nw + (t >> 6)] |= bit;
        }
    }
    // the most used offsets among the edges the first masks leave
    std::vector<double> by_off(127, 0.0);  // offset d at d + 63
    for (int32_t s = 0; s < n; ++s) {
        if (s == start && !g.fmask.empty()) continue;
        for (int32_t k = f.srp[s]; k < f.srp[s + 1]; ++k) {
            const int32_t d = f.st[k] - s;
            if (d != 0 && d >= -63 && d <= 63) by_off[static_cast<size_t>(d + 63)] += weight(f.ss[k]);
        }
    }
    std::vector<int32_t> order;
    for (int32_t d = -63; d <= 63; ++d)
        if (d != 0 && by_off[static_cast<size_t>(d + 63)] >= keep) order.push_back(d);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return by_off[static_cast<size_t>(x + 63)] > by_off[static_cast<size_t>(y + 63)];
    });
    g.ndiag = static_cast<int32_t>(std::min<size_t>(order.size(), kMaxDiag));
    for (int32_t k = 0; k < g.ndiag; ++k) g.doff[k] = order[static_cast<size_t>(k)];
    if (!g.ndiag) return;
    const size_t nd = static_cast<size_t>(g.ndiag);
    g.dmask.assign(static_cast<size_t>(kAlphabet) * nd * nw, 0ull);
    for (int32_t s = 0; s < n; ++s) {
        if (s == start && !g.fmask.empty()) continue;
        for (int32_t k = f.srp[s]; k < f.srp[s + 1]; ++k) {
            const int32_t d = f.st[k] - s;
            for (int32_t q = 0; q < g.ndiag; ++q) {
                if (g.doff[q] != d) continue;
                const int32_t t = f.st[k];  // indexed by target
                const uint64_t bit = 1ull << (t & 63);
                const int c0 = f.ss[k] == kAnySymbol ? 0 : f.ss[k];
                const int c1 = f.ss[k] == kAnySymbol ? kAlphabet : f.ss[k] + 1;
                for (int c = c0; c < c1; ++c)
                    g.dmask[(static_cast<size_t>(c) * nd + static_cast<size_t>(q)) * nw +
                            static_cast<size_t>(t >> 6)] |= bit;
            }
        }
    }
}

}  // namespace

GeneralHost build_general(const NfaInput &a, DiagMode mode) {
    GeneralHost g;
    FoldedHost f = fold_epsilon(a);
    if (!f.ok) {
        g.why = f.why;
        return g;
    }
    const int32_t n = a.n;
    std::vector<int32_t> no_eps(static_cast<size_t>(n) + 1, 0);
    NfaInput b{n, a.start, a.accept, f.srp.data(), f.st.data(), f.ss.data(), no_eps.data(),
               nullptr};
    g.rows = build_worklist(b);
    g.rows.start_list = f.start_set;
    g.rows.start_accepts = f.start_accepts;
    g.start_bits.assign(static_cast<size_t>(g.rows.nwords), 0ull);
    for (int32_t s : f.start_set) g.start_bits[s >> 6] |= 1ull << (s & 63);
    if (n <= 65536) {
        build_gather_masks(g, f, a.start, n, mode);
        // counting sort of the folded edges left to the lists by label (ANY last)
        auto listed = [&](int32_t s, int32_t k) { return !mask_covers(g, a.start, s, f.st[k]); };
        g.boff.assign(kAnySymbol + 2, 0);
        for (int32_t s = 0; s < n; ++s)
            for (int32_t k = f.srp[s]; k < f.srp[s + 1]; ++k)
                if (listed(s, k)) g.boff[f.ss[k] + 1]++;
        for (int c = 0; c <= kAnySymbol; ++c) g.boff[c + 1] += g.boff[c];
        std::vector<int32_t> cur(g.boff.begin(), g.boff.end() - 1);
        g.bedge.resize(static_cast<size_t>(g.boff[kAnySymbol + 1]));
        for (int32_t s = 0; s < n; ++s)
            for (int32_t k = f.srp[s]; k < f.srp[s + 1]; ++k)
                if (listed(s, k))
                    g.bedge[cur[f.ss[k]]++] = (static_cast<uint32_t>(s) << 16) |
                                              static_cast<uint32_t>(f.st[k]);
        g.gather_ok = true;
    }
    g.ok = true;
    return g;
}

DfaHost build_dfa(int32_t n, int32_t start, const uint8_t *accept, const int32_t *trans) {
    if (n < 1) bad("num_states must be >= 1");
    if (start < 0 || start >= n) bad("start_state out of range");
    if (!accept || !trans) bad("null array");
    if (static_cast<int64_t>(n) * kAlphabet * 4 >= (1ll << 31)) bad("DFA table too large");
    const size_t total = static_cast<size_t>(n) * kAlphabet;
    for (size_t k = 0; k < total; ++k)
        if (trans[k] < 0 || trans[k] >= n)
            bad("trans entry out of range at index " + std::to_string(k));
    DfaHost h;
    h.esz = n <= 256 ? 1u : (n <= 0x7fff ? 2u : 4u);
    h.wide = h.esz == 4;
    const uint32_t esz = h.esz;
    // non-accepting states first, accepting states last (original order kept in each)
    std::vector<uint32_t> id(static_cast<size_t>(n));
    uint32_t next = 0;
    for (int32_t s = 0; s < n; ++s)
        if (!accept[s]) id[s] = next++;
    const uint32_t first_accepting = next;
    for (int32_t s = 0; s < n; ++s)
        if (accept[s]) id[s] = next++;
    h.accept_from = first_accepting * esz;
    h.start_entry = id[start] * esz;
    const size_t colen = esz == 1 ? 256 : static_cast<size_t>(n);  // entries per column
    h.col_bytes = static_cast<uint32_t>(colen) * esz;
    h.start_accepts = accept[start] != 0;
    const size_t cells = colen * kAlphabet;
    if (esz == 1) h.t8.assign(cells, 0);
    else if (esz == 2) h.t16.resize(cells);
    else h.t32.resize(cells);
    for (int32_t s = 0; s < n; ++s) {
        const uint32_t row = id[s];
        for (int c = 0; c < kAlphabet; ++c) {
            const uint32_t e = id[trans[static_cast<size_t>(s) * kAlphabet + c]] * esz;
            const size_t at = static_cast<size_t>(c) * colen + row;
            if (esz == 1) h.t8[at] = static_cast<uint8_t>(e);
            else if (esz == 2) h.t16[at] = static_cast<uint16_t>(e);
            else h.t32[at] = e;
        }
    }
    return h;
}

}  // namespace gfsm
