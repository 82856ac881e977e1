// Host-side check: the per-photon split of Philox4x32-10 used by the
// propagation kernel (ntp::PhiloxRun) against the plain ntp::philox on
// random keys, counters and block indices (both functions are host-callable).
#include "ntprop_device.h"
#include <cstdio>
#include <random>
int main() {
    std::mt19937_64 g(1);
    long bad = 0, n = 0;
    for (int t = 0; t < 200000; ++t) {
        const uint32_t k0 = g(), k1 = g(), lo = g(), hi = (t % 3) ? 0u : (uint32_t)g(), src = g();
        ntp::PhiloxRun run; run.init(k0, k1, lo, hi, src);
        for (uint32_t b : {1u, 2u, 3u, 17u, (uint32_t)g()}) {
            const uint4 a = run.block(b), e = ntp::philox(make_uint4(b, lo, hi, src), k0, k1);
            ++n; bad += (a.x != e.x) | (a.y != e.y) | (a.z != e.z) | (a.w != e.w);
        }
    }
    printf("checked %ld blocks, mismatches %ld\n", n, bad);
    return bad != 0;
}
