// Writes Philox4x32-10 blocks produced by rocRAND's open-source device/host
// header (rocrand_philox4x32_10.h) for a set of keys and counters, as a
// known-answer fixture for the project's own Philox implementation.
//   hipcc -x hip --offload-arch=gfx950 -O1 -o /tmp/philox_gen tests/golden/philox_rocrand_gen.cpp
//   /tmp/philox_gen > tests/golden/philox_rocrand.txt
// Line format: seed subsequence block  w0 w1 w2 w3   (all hex)
// rocrand_init(seed, subsequence, 4*block) positions the counter at
// (block_lo, block_hi, subsequence_lo, subsequence_hi) with key (seed_lo, seed_hi).
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
int main() {
  const unsigned long long seeds[] = {0ULL, 1ULL, 0x0123456789abcdefULL, 0xffffffffffffffffULL};
  const unsigned long long subs[] = {0ULL, 7ULL, 0x100000003ULL};
  const unsigned long long blocks[] = {0ULL, 1ULL, 12345ULL, 0x1fffffffULL};
  for (auto s : seeds)
    for (auto q : subs)
      for (auto b : blocks) {
        rocrand_state_philox4x32_10 st;
        rocrand_init(s, q, 4 * b, &st);
        uint4 v = rocrand4(&st);
        printf("%llx %llx %llx %08x %08x %08x %08x\n", s, q, b, v.x, v.y, v.z, v.w);
      }
  return 0;
}
