// Host check: the device permutation stream (perm_rng.hpp) against libstdc++'s
// std::mt19937 + std::shuffle on the same seeds, for many n (odd and even).
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../eqtl_amd/csrc/perm_rng.hpp"

int main() {
  uint32_t st[eqtl_amd::kMtN];
  long checked = 0;
  const char* genes[] = {"g0", "g1", "ENSG00000141510", "gene_with_long_name_x"};
  for (int n : {1, 2, 3, 4, 5, 10, 99, 100, 257, 1000, 1001, 2000, 4999, 5000}) {
    for (const char* gene : genes) {
      for (int seed : {-1, 0, 42, 123456}) {
        for (int b = 0; b < 40; b += 7) {
          const uint32_t s = eqtl_amd::perm_seed(seed, b, eqtl_amd::fnv1a32(gene));
          std::mt19937 ref(s);
          std::vector<int> a(n);
          std::iota(a.begin(), a.end(), 0);
          std::shuffle(a.begin(), a.end(), ref);
          eqtl_amd::Mt19937 g{st, 0};
          g.seed(s);
          std::vector<int> d(n);
          eqtl_amd::shuffle_iota(g, d.data(), n);
          if (a != d) {
            std::printf("MISMATCH n=%d gene=%s seed=%d b=%d\n", n, gene, seed, b);
            return 1;
          }
          ++checked;
        }
      }
    }
  }
  std::printf("perm streams identical: %ld cases\n", checked);
  return 0;
}
