// Writes rocRAND Philox4x32-10 uniforms (rocrand_init(seed, subsequence, offset)
// + rocrand_uniform, run on the host) as text: seed subseq offset float_bits.
// Built and run by make_philox_vectors.py; the output is the committed fixture.
#include <cstdio>
#include <cstring>
#include <cstdint>
#include <rocrand/rocrand_kernel.h>

int main() {
    const uint32_t seeds[] = {0u, 1u, 42u, 54321u, 0xFFFFFFFFu};
    const uint32_t subs[] = {0u, 1u, 7u, 262143u, 4294967295u, 2073599u};
    const uint32_t offs[] = {0u, 1u, 2u, 3u, 4u, 5u, 13u, 63u, 1000u, 65537u};
    for (uint32_t s : seeds)
        for (uint32_t q : subs)
            for (uint32_t o : offs) {
                rocrand_state_philox4x32_10 st;
                rocrand_init(s, q, o, &st);
                float u = rocrand_uniform(&st);
                uint32_t bits;
                std::memcpy(&bits, &u, 4);
                std::printf("%u %u %u %u\n", s, q, o, bits);
            }
    return 0;
}
