// Check of the staging copy pool (csrc/copy_pool.hpp), built with ThreadSanitizer by
// tests/test_native_host.py (CPU suite; no GPU).
//
// Many copies in quick succession from one caller, of sizes around the pool's part size
// (below two parts the caller copies alone), each into a buffer with guard bytes on both
// sides: every destination byte equals its source and no guard byte changes. The next copy
// reuses the same destination with other contents, so a helper that wrote the previous copy's
// bytes late would leave wrong bytes; ThreadSanitizer reports unsynchronised access between
// the caller and the helpers. Pools of 0, 1, 3 and 7 helpers; a pool destroyed right after it
// is made, and one right after a copy.
// Exit status 0 and one summary line on success; the first mismatch is printed and exits 1.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "copy_pool.hpp"

namespace {

int fail(const char *what, size_t helpers, size_t n) {
    std::printf("pool_check: %s (helpers %zu, copy of %zu bytes)\n", what, helpers, n);
    return 1;
}

}  // namespace

int main() {
    constexpr size_t kMiB = size_t(1) << 20;
    constexpr size_t kGuard = 4096;
    constexpr uint8_t kGuardByte = 0xa5;
    const size_t sizes[] = {0,        1,           kMiB - 1,     kMiB,          2 * kMiB - 1,
                            2 * kMiB, 2 * kMiB + 1, 3 * kMiB + 7, 8 * kMiB,      8 * kMiB + 12345,
                            17 * kMiB + 3};
    const size_t biggest = 17 * kMiB + 3;
    // one random source; each copy starts at another offset into it, so consecutive copies
    // into the same destination carry different bytes
    std::vector<uint8_t> src(biggest + 4096), dst(biggest + 2 * kGuard), guard(kGuard, kGuardByte);
    std::mt19937_64 rng(7);
    for (size_t i = 0; i + 8 <= src.size(); i += 8) {
        const uint64_t w = rng();
        std::memcpy(src.data() + i, &w, 8);
    }
    int64_t copies = 0;
    for (size_t helpers : {size_t(0), size_t(1), size_t(3), size_t(7)}) {
        gfsm::CopyPool pool(helpers);
        for (int round = 0; round < 3; ++round) {
            for (size_t n : sizes) {
                const size_t shift = rng() % 4096;  // unaligned source
                std::memset(dst.data(), kGuardByte, dst.size());
                pool.copy(dst.data() + kGuard, src.data() + shift, n);
                ++copies;
                if (n && std::memcmp(dst.data() + kGuard, src.data() + shift, n) != 0)
                    return fail("wrong bytes", helpers, n);
                if (std::memcmp(dst.data(), guard.data(), kGuard) != 0 ||
                    std::memcmp(dst.data() + kGuard + n, guard.data(), kGuard) != 0)
                    return fail("guard byte written", helpers, n);
            }
        }
    }
    { gfsm::CopyPool idle(5); }  // made and destroyed without a copy
    {
        gfsm::CopyPool once(3);
        once.copy(dst.data(), src.data(), 4 * kMiB);
    }  // destroyed right after a copy
    std::printf("pool_check ok: %lld copies\n", static_cast<long long>(copies));
    return 0;
}
