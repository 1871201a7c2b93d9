// Prints Philox4x32-10 blocks from rocRAND's header-only engine (host path)
// for fixed (seed, subsequence) pairs: counter = {0 + k, 0, subseq_lo, subseq_hi},
// key = {seed_lo, seed_hi}.  Used to write tests/golden/philox_rocrand.json.
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

int main() {
  const unsigned long long seeds[] = {0ull, 42ull, 0xdeadbeefcafef00dull, 0xffffffffffffffffull};
  const unsigned long long subs[] = {0ull, 1ull, 7ull, 0x123456789ull};
  for (unsigned long long seed : seeds)
    for (unsigned long long sub : subs) {
      rocrand_state_philox4x32_10 st;
      rocrand_init(seed, sub, 0, &st);
      for (int b = 0; b < 2; b++) {
        unsigned v[4];
        for (int k = 0; k < 4; k++) v[k] = rocrand(&st);
        printf("%llu %llu %d %u %u %u %u\n", seed, sub, b, v[0], v[1], v[2], v[3]);
      }
    }
  return 0;
}
