// order_check.cpp - host-side check of the leader ordering of the cycle-following 010
// (lm::order_010_leaders, an internal function of the library): for many tile shapes and
// element sizes, the list must be padded to whole groups of 64 and the walks from its starts
// must cover every slot of [1, a*B - 1) that is not a fixed point exactly once. The count of
// cycles it is given comes from a plain walk here.
#include <cstdint>
#include <cstdio>
#include <vector>

namespace lm {
std::vector<uint32_t> order_010_leaders(uint32_t a, uint32_t B, size_t es, uint64_t cycles);
}

int main() {
  int fails = 0, ordered = 0, shapes = 0;
  const size_t sizes[] = {1, 2, 4, 8, 16};
  for (uint32_t a = 2; a <= 70; a += 3)
    for (uint32_t B = 2; B <= 70; B += 5)
      for (size_t es : sizes) {
        const uint32_t n = a * B, last = n - 1;
        auto next = [&](uint32_t s) { return (s % B) * a + s / B; };
        std::vector<uint8_t> seen(n, 0);
        uint64_t cycles = 0, moved = 0;
        uint32_t longest = 0;
        for (uint32_t s = 1; s < last; s++) {
          if (seen[s]) continue;
          uint32_t t = s, len = 0;
          do {
            seen[t] = 1;
            len++;
            t = next(t);
          } while (t != s);
          if (len > 1) {
            cycles++;
            moved += len;
          }
          longest = len > longest ? len : longest;
        }
        shapes++;
        const std::vector<uint32_t> v = lm::order_010_leaders(a, B, es, cycles);
        if (v.empty()) {
          // only tiles with a cycle longer than 64 slots, or none to move, are left unordered
          if (longest <= 64 && cycles > 0) {
            std::printf("a=%u B=%u es=%zu: not ordered\n", a, B, es);
            fails++;
          }
          continue;
        }
        ordered++;
        if (v.size() % 64 != 0) {
          std::printf("a=%u B=%u es=%zu: %zu entries\n", a, B, es, v.size());
          fails++;
        }
        std::vector<uint8_t> hit(n, 0);
        uint64_t covered = 0, starts = 0;
        bool bad = false;
        for (uint32_t s : v) {
          if (s == 0xffffffffu) continue;
          starts++;
          uint32_t t = s;
          do {
            if (t == 0 || t >= last || hit[t]) bad = true;
            else hit[t] = 1;
            covered++;
            t = next(t);
          } while (t != s && !bad);
        }
        if (bad || covered != moved || starts != cycles) {
          std::printf("a=%u B=%u es=%zu: starts %llu of %llu cycles, %llu of %llu slots\n", a, B,
                      es, (unsigned long long)starts, (unsigned long long)cycles,
                      (unsigned long long)covered, (unsigned long long)moved);
          fails++;
        }
      }
  std::printf("shapes=%d ordered=%d fails=%d\n", shapes, ordered, fails);
  return fails ? 1 : 0;
}
