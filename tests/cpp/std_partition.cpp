// std::partition of (code, id) pairs by (code & mask) == flag, with the C++ library's
// own algorithm: the independent check of the oracle's partition scheme (the
// reference's Partitioning::partition, detail/partitioning.h:215-241, and the
// std::partition it calls for non-refinable boxes, ot_static_linear_core.h:455-456).
// stdin: n mask flag, then n lines "code id"; stdout: the front count, then the ids.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <list>
#include <utility>
#include <vector>

int main() {
  unsigned long long n = 0, mask = 0, flag = 0;
  if (std::scanf("%llu %llu %llu", &n, &mask, &flag) != 3) return 1;
  // a std::list has bidirectional iterators: the two-pointer __partition overload
  std::list<std::pair<uint64_t, uint32_t>> v;
  for (unsigned long long i = 0; i < n; ++i) {
    unsigned long long c = 0;
    unsigned id = 0;
    if (std::scanf("%llu %u", &c, &id) != 2) return 1;
    v.emplace_back(c, id);
  }
  auto mid = std::partition(v.begin(), v.end(), [&](const auto &x) { return (x.first & mask) == flag; });
  std::printf("%ld\n", (long)std::distance(v.begin(), mid));
  for (const auto &x : v) std::printf("%u\n", x.second);
  return 0;
}
