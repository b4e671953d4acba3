// CPU check of the per-film sweep tile choice (sweep_film_tile in
// tdgl_launch.hpp) on every film size from 1 to 400 nodes across and along:
// the tile fits the sweep block (16 <= W <= 64, W R <= 256 lanes, at most two
// staged cells per lane, (W + 2)(R + 2) within the shared memory's
// kSweepCells), the LDS cell c0 that every lane of the block reads stays
// inside the staged cells (also for the lanes past W x R, which compute
// nothing), and no other admissible tile has fewer lanes-with-cost per plane
// (the choice is the argmin it claims, ties to the wider tile).  Host code
// only: hipcc compiles the header's inline functions for the CPU.
#include <stdio.h>

#include "tdgl_launch.hpp"

using namespace tdgl;

int main() {
  long long bad = 0, films = 0;
  for (long long nx = 1; nx <= 400; ++nx)
    for (long long ny = 1; ny <= 400; ++ny) {
      int w = 0, h = 0;
      sweep_film_tile(nx, ny, w, h);
      ++films;
      const bool fits = w >= 16 && w <= 64 && h >= 1 && w * h <= 256 && (w + 2) * (h + 2) <= kSweepCells &&
                        (w + 2) * (h + 2) <= 2 * 256;
      int cmax = 0;
      for (int l = 0; l < 256; ++l) {
        const int c0 = (l / w + 1) * (w + 2) + (l % w) + 1;
        if (c0 > cmax) cmax = c0;
      }
      const double cost = (double)(((nx + w - 1) / w) * ((ny + h - 1) / h)) * (23.0 + 64.0 / w);
      bool best = true;
      for (int w2 = 16; w2 <= 64 && best; ++w2)
        for (int h2 = 1; w2 * h2 <= 256; ++h2) {
          if ((w2 + 2) * (h2 + 2) > kSweepCells) continue;
          const double c2 = (double)(((nx + w2 - 1) / w2) * ((ny + h2 - 1) / h2)) * (23.0 + 64.0 / w2);
          if (c2 < cost || (c2 == cost && w2 > w)) best = false;
        }
      if (!fits || cmax >= kSweepCells || !best) {
        if (bad < 10)
          printf("film %lld x %lld: tile %d x %d fits %d c0max %d best %d\n", nx, ny, w, h, (int)fits, cmax,
                 (int)best);
        ++bad;
      }
    }
  if (bad) {
    printf("FAIL %lld of %lld films\n", bad, films);
    return 1;
  }
  printf("OK %lld films\n", films);
  return 0;
}
