// Writes Philox4x32-10 vectors computed by rocRAND's host-callable engine
// (an independent implementation of the same published generator) as JSON:
// for each (seed, counter) the 4 words of philox(ctr = (lo, hi, 0, 0), key = seed).
// Build and run: tests/golden/make_golden.py.
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

int main() {
    const unsigned long long seeds[] = {0ull, 1ull, 12345ull, 0xdeadbeefcafef00dull};
    const unsigned long long counters[] = {0ull, 1ull, 2ull, 1000003ull, 0xffffffffull,
                                           0x100000000ull, 0x123456789abcull};
    std::printf("[\n");
    bool first = true;
    for (unsigned long long s : seeds) {
        for (unsigned long long c : counters) {
            // offset 4*c: counter (lo(c), hi(c), 0, 0); next4() returns that block
            rocrand_device::philox4x32_10_engine eng(s, 0ull, 4ull * c);
            const uint4 r = eng.next4();
            std::printf("%s  {\"seed\": %llu, \"counter\": %llu, \"out\": [%u, %u, %u, %u]}",
                        first ? "" : ",\n", s, c, r.x, r.y, r.z, r.w);
            first = false;
        }
    }
    std::printf("\n]\n");
    return 0;
}
