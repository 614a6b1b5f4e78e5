// Host paths of the C ABI (csrc/capi.hip) under AddressSanitizer and UBSan, without a GPU:
// built by tests/test_native_host.py from capi.hip compiled as host C++, the layout builders,
// and the library's kernel objects (which are linked but never launched here).
//
// Checked: the ABI version; gfsm_nfa_analyze on general, chain-structured and fold-capped
// automata (the host layouts, the packer's size accounting and the reported properties);
// argument validation and status codes of every entry point that can refuse its arguments
// before touching a device; the thread-local last-error message.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gpufsm_hip.h"

namespace {

[[noreturn]] void fail(const char *what, long a = 0) {
    std::fprintf(stderr, "capi_check FAILED: %s (%ld)\n", what, a);
    std::exit(1);
}

int g_checks = 0;

void expect(int rc, int want, const char *what) {
    if (rc != want) {
        std::fprintf(stderr, "capi_check FAILED: %s: status %d, want %d (%s)\n", what, rc, want,
                     gfsm_last_error());
        std::exit(1);
    }
    ++g_checks;
}

struct Csr {
    int32_t n = 0, start = 0;
    std::vector<uint8_t> acc;
    std::vector<int32_t> srp{0}, st, ss, erp{0}, et;
    void row(const std::vector<std::pair<int32_t, int32_t>> &edges,
             const std::vector<int32_t> &eps) {
        for (auto [t, l] : edges) {
            st.push_back(t);
            ss.push_back(l);
        }
        srp.push_back(static_cast<int32_t>(st.size()));
        for (int32_t t : eps) et.push_back(t);
        erp.push_back(static_cast<int32_t>(et.size()));
    }
    int analyze(gfsm_nfa_props *p) {
        st.reserve(st.size() + 1);
        ss.reserve(ss.size() + 1);
        et.reserve(et.size() + 1);
        return gfsm_nfa_analyze(n, start, acc.data(), srp.data(), st.data(), ss.data(),
                                erp.data(), et.data(), p);
    }
};

// A general automaton: an epsilon cycle through every state and labelled edges.
Csr general(int32_t n) {
    Csr c;
    c.n = n;
    c.acc.assign(n, 0);
    c.acc[n - 1] = 1;
    for (int32_t s = 0; s < n; ++s)
        c.row({{(s * 7 + 3) % n, 'a' + s % 5}, {(s + 11) % n, 256}}, {(s + 1) % (n < 4 ? n : 4)});
    return c;
}

// A pattern union: start with an ANY self-loop and chains of 3 positions.
Csr union_of(int32_t patterns) {
    Csr c;
    c.n = 1 + 3 * patterns;
    c.acc.assign(c.n, 0);
    std::vector<std::pair<int32_t, int32_t>> start{{0, 256}};
    for (int32_t p = 0; p < patterns; ++p) start.push_back({1 + 3 * p, (p * 13) % 256});
    c.row(start, {});
    for (int32_t p = 0; p < patterns; ++p) {
        c.acc[3 + 3 * p] = 1;
        c.row({{2 + 3 * p, (p * 5) % 256}}, {});
        c.row({{3 + 3 * p, (p * 3 + 1) % 256}}, {});
        c.row({}, {});
    }
    return c;
}

}  // namespace

int main() {
    if (gfsm_abi_version() != GFSM_ABI_VERSION) fail("abi version", gfsm_abi_version());
    ++g_checks;

    gfsm_nfa_props p;
    // general automata: not chain-structured, folded layout present
    for (int32_t n : {1, 5, 64, 65, 300, 5000}) {
        Csr c = general(n);
        expect(c.analyze(&p), GFSM_OK, "analyze general");
        if (p.chain_ok || !p.general_ok || p.num_states != n || p.device_bytes <= 0)
            fail("general props", n);
        if (p.num_eps_edges != static_cast<int64_t>(c.et.size())) fail("eps count", n);
    }
    // pattern unions: chain-structured; the folded layout only up to 512 states
    for (int32_t k : {10, 170, 171, 2000}) {
        Csr c = union_of(k);
        expect(c.analyze(&p), GFSM_OK, "analyze union");
        if (!p.chain_ok || !p.start_always || p.num_tiles < 1 || p.max_chain != 3)
            fail("union props", k);
        if (p.general_ok != (c.n <= 512)) fail("union general layout", c.n);
    }
    // a closure too large to fold: 3000-state epsilon cycle
    {
        Csr c;
        c.n = 3000;
        c.acc.assign(c.n, 0);
        for (int32_t s = 0; s < c.n; ++s) c.row({{s, 'x'}}, {(s + 1) % c.n});
        expect(c.analyze(&p), GFSM_OK, "analyze fold cap");
        if (p.general_ok || p.chain_ok) fail("fold cap props");
    }
    // an accepting start closure
    {
        Csr c;
        c.n = 2;
        c.acc = {0, 1};
        c.row({}, {1});
        c.row({}, {});
        expect(c.analyze(&p), GFSM_OK, "analyze accepting start");
        if (!p.start_accepts) fail("start_accepts");
    }

    // validation: malformed automata and bad arguments, before any device work
    {
        Csr c = general(8);
        c.st[3] = 8;  // target out of range
        expect(c.analyze(&p), GFSM_EINVAL, "target out of range");
        if (std::strstr(gfsm_last_error(), "sym_targets") == nullptr) fail("error message");
        c = general(8);
        c.ss[0] = 300;
        expect(c.analyze(&p), GFSM_EINVAL, "label past 256");
        c = general(8);
        c.start = 9;
        expect(c.analyze(&p), GFSM_EINVAL, "start out of range");
        c = general(8);
        expect(gfsm_nfa_analyze(8, 0, c.acc.data(), c.srp.data(), c.st.data(), c.ss.data(),
                                c.erp.data(), c.et.data(), nullptr),
               GFSM_EINVAL, "null props");
        expect(gfsm_nfa_analyze(0, 0, c.acc.data(), c.srp.data(), c.st.data(), c.ss.data(),
                                c.erp.data(), c.et.data(), &p),
               GFSM_EINVAL, "zero states");
    }
    int32_t tech = 0;
    int64_t bytes = 0;
    int32_t flags[1], lens[1];
    const int64_t off[2] = {0, 0};
    gfsm_tuning bad_tuning = {0, 0, 0, -1};
    expect(gfsm_nfa_create(1, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr),
           GFSM_EINVAL, "create null out");
    expect(gfsm_dfa_create(1, 0, nullptr, nullptr, nullptr), GFSM_EINVAL, "dfa create null out");
    expect(gfsm_nfa_run_host(nullptr, 0, nullptr, off, 1, flags, lens, nullptr), GFSM_EINVAL,
           "run_host null handle");
    expect(gfsm_dfa_run_host(nullptr, 0, nullptr, off, 1, flags, lens, nullptr), GFSM_EINVAL,
           "dfa run_host null handle");
    expect(gfsm_nfa_run_device(nullptr, 0, nullptr, off, 1, 0, flags, lens, nullptr), GFSM_EINVAL,
           "run_device null handle");
    expect(gfsm_nfa_resolve(nullptr, 0, nullptr, 1, 1, &tech), GFSM_EINVAL, "resolve null");
    expect(gfsm_dfa_resolve(nullptr, 0, nullptr, 1, 1, &tech), GFSM_EINVAL, "dfa resolve null");
    expect(gfsm_nfa_workspace_size(nullptr, 0, 1, 1, nullptr, &bytes), GFSM_EINVAL,
           "workspace null");
    expect(gfsm_dfa_workspace_size(nullptr, 0, 1, 1, &bad_tuning, &bytes), GFSM_EINVAL,
           "dfa workspace null");
    expect(gfsm_nfa_run_device_ws(nullptr, 0, nullptr, nullptr, off, 1, 0, flags, lens, nullptr,
                                  0, nullptr),
           GFSM_EINVAL, "run_device_ws null");
    expect(gfsm_nfa_destroy(nullptr), GFSM_OK, "destroy null is a no-op");
    expect(gfsm_dfa_destroy(nullptr), GFSM_OK, "dfa destroy null is a no-op");
    char arch[8];
    expect(gfsm_device_arch(0, nullptr, 8), GFSM_EINVAL, "arch null buffer");
    expect(gfsm_device_arch(0, arch, 0), GFSM_EINVAL, "arch zero length");

    // the last-error message is per thread
    expect(gfsm_nfa_resolve(nullptr, 0, nullptr, 1, 1, &tech), GFSM_EINVAL, "resolve null again");
    const std::string here = gfsm_last_error();
    std::string there = "unset";
    std::thread th([&] { there = gfsm_last_error(); });
    th.join();
    if (here.empty() || !there.empty()) fail("thread-local error message");

    std::printf("capi_check ok: %d checks\n", g_checks);
    return 0;
}
