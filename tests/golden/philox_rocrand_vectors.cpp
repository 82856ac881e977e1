// Prints Philox4x32-10 outputs of rocRAND's header-only host implementation
// for counter 0 and several keys; the printed words are the "rocrand" entries
// of philox4x32_10_kat.json.  Build: hipcc -O1 philox_rocrand_vectors.cpp
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
int main() {
    const unsigned long long seeds[3] = {0ull, 0xffffffffffffffffull, 0x299f31d0a4093822ull};
    for (unsigned long long s : seeds) {
        rocrand_state_philox4x32_10 st;
        rocrand_init(s, 0, 0, &st);  // key = seed, counter = 0
        unsigned a = rocrand(&st), b = rocrand(&st), c = rocrand(&st), d = rocrand(&st);
        printf("%016llx %08x %08x %08x %08x\n", s, a, b, c, d);
    }
    return 0;
}
