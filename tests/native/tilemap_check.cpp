// Host-side check of the persistent-grid tile schedule (caps_amd/csrc/tilemap.h):
// for every grid persistent_grid can return, xcd_tile is a permutation of
// [0, G), and the rounds of the grid cover every tile id exactly once.
#include <cstdio>
#include <vector>

#include "../../caps_amd/csrc/tilemap.h"

int main() {
  using namespace caps;
  const int totals[] = {1, 7, 8, 9, 63, 64, 100, 255, 256, 257, 1000, 4096, 43904};
  long long checked = 0;
  for (int cus = 1; cus <= 320; cus++)
    for (int bpc = 1; bpc <= 2; bpc++)
      for (int total : totals) {
        const int G = persistent_grid(bpc, cus, total);
        if (G <= 0 || G % NXCD != 0) {
          printf("FAIL grid cus=%d bpc=%d total=%d G=%d\n", cus, bpc, total, G);
          return 1;
        }
        std::vector<int> seen((size_t)G, 0);
        for (int b = 0; b < G; b++) {
          const int t = xcd_tile(b, G);
          if (t < 0 || t >= G || seen[(size_t)t]++) {
            printf("FAIL permutation cus=%d bpc=%d total=%d G=%d b=%d\n", cus, bpc, total, G, b);
            return 1;
          }
        }
        std::vector<int> hit((size_t)total, 0);
        for (int round = 0; round * G < total; round++)
          for (int b = 0; b < G; b++) {
            const int t = round * G + xcd_tile(b, G);
            if (t < total) hit[(size_t)t]++;
          }
        for (int t = 0; t < total; t++)
          if (hit[(size_t)t] != 1) {
            printf("FAIL cover cus=%d bpc=%d total=%d tile=%d hits=%d\n", cus, bpc, total, t,
                   hit[(size_t)t]);
            return 1;
          }
        checked++;
      }
  printf("OK %lld grids\n", checked);
  return 0;
}
