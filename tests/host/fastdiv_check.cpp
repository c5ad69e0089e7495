// Host-side check of the library's division by invariant integers (csrc/common.h):
// quotients must equal the hardware divide for edge and random operands.
#include "../../libmarshal_amd/csrc/common.h"

#include <cstdio>
#include <random>
#include <vector>

int main() {
  std::mt19937_64 rng(12345);
  long fails = 0;
  std::vector<uint32_t> d32;
  for (uint32_t d = 1; d <= 3000; d++) d32.push_back(d);
  for (int s = 0; s < 32; s++) { d32.push_back(1u << s); d32.push_back((1u << s) + 1); d32.push_back((1u << s) - 1 ? (1u << s) - 1 : 1); }
  d32.push_back(0xffffffffu); d32.push_back(0xfffffffeu); d32.push_back(0x80000001u);
  for (int k = 0; k < 2000; k++) { uint32_t d = uint32_t(rng()); d32.push_back(d ? d : 1); }
  for (uint32_t d : d32) {
    lm::FastDiv32 f(d);
    uint32_t edge[] = {0, 1, d - 1, d, d + 1, 2 * d - 1, 2 * d, 0xffffffffu, 0xfffffffeu, 0x80000000u};
    for (uint32_t n : edge) if (f.div(n) != n / d) { if (fails++ < 10) printf("32: %u/%u\n", n, d); }
    for (int k = 0; k < 300; k++) { uint32_t n = uint32_t(rng()); if (f.div(n) != n / d) { if (fails++ < 10) printf("32: %u/%u\n", n, d); } }
  }
  std::vector<uint64_t> d64;
  for (uint64_t d = 1; d <= 2000; d++) d64.push_back(d);
  for (int s = 0; s < 64; s++) { d64.push_back(1ull << s); d64.push_back((1ull << s) + 1); if (s) d64.push_back((1ull << s) - 1); }
  d64.push_back(~0ull); d64.push_back(250007ull * 199999ull - 1); d64.push_back(98304ull * 131071ull - 1);
  for (int k = 0; k < 2000; k++) { uint64_t d = rng() >> (rng() % 64); d64.push_back(d ? d : 1); }
  for (uint64_t d : d64) {
    lm::FastDiv64 f(d);
    uint64_t edge[] = {0, 1, d - 1, d, d + 1, 2 * d - 1, 2 * d, ~0ull, ~0ull - 1, 1ull << 63, 50000000000ull};
    for (uint64_t n : edge) if (f.div(n) != n / d) { if (fails++ < 10) printf("64: %llu/%llu\n", (unsigned long long)n, (unsigned long long)d); }
    for (int k = 0; k < 200; k++) { uint64_t n = rng() >> (rng() % 64); if (f.div(n) != n / d) { if (fails++ < 10) printf("64: %llu/%llu\n", (unsigned long long)n, (unsigned long long)d); } }
  }
  printf("fails=%ld\n", fails);
  return fails ? 1 : 0;
}
