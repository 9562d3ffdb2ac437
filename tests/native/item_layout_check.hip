// Host check of the step kernels' bulk item layout (set_item_layout, item_cell and
// last_item_of_row in windtunnel_vis_amd/csrc/lbm.hpp, compiled from the same header). For every grid
// given as nx,ny,nz,cpt arguments: every cell is worked by exactly one lane of one item; every lane
// that works no cell loads inside its own row (its chunk's first x); an item's row is the same for
// all of its lanes; last_item_of_row names the item holding each row's last cell (the role table
// places wall blocks by it). Prints "ok <grids> <items>" or the first failure (exit 1).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../windtunnel_vis_amd/csrc/lbm.hpp"

using namespace wt;

static int fail(const char* what, long nx, long ny, long nz, int cpt, long item, int lane) {
  std::printf("FAIL %s grid %ldx%ldx%ld cpt %d item %ld lane %d\n", what, nx, ny, nz, cpt, item, lane);
  return 1;
}

int main(int argc, char** argv) {
  long items_total = 0;
  int grids = 0;
  for (int a = 1; a < argc; a++) {
    long nx, ny, nz;
    int cpt;
    if (std::sscanf(argv[a], "%ld,%ld,%ld,%d", &nx, &ny, &nz, &cpt) != 4) return fail("bad argument", 0, 0, 0, 0, a, 0);
    StepArgs A{};
    set_item_layout(A, nx, ny, nz, cpt);
    const long rows = ny * nz;
    std::vector<unsigned char> seen((size_t)(nx * rows), 0);
    std::vector<long> owner((size_t)rows, -1);  // the item holding each row's last cell
    for (long item = 0; item < (long)A.nwaves; item++) {
      uint32_t row0 = 0;
      for (int lane = 0; lane < 64; lane++) {
        const ItemCell c = item_cell(A, (uint32_t)item, (uint32_t)lane, cpt);
        if (lane == 0) row0 = c.row;
        if (c.row != row0) return fail("row differs between lanes", nx, ny, nz, cpt, item, lane);
        const long row = (long)c.row;
        if (row >= rows) return fail("row out of range", nx, ny, nz, cpt, item, lane);
        if (c.xs < 0 || c.xs >= nx) return fail("chunk start outside the row", nx, ny, nz, cpt, item, lane);
        if (c.x0 < c.xs) return fail("lane before its chunk", nx, ny, nz, cpt, item, lane);
        for (int k = 0; k < cpt; k++) {
          const long x = c.x0 + k;
          if (x >= nx) continue;  // past the row end: works no cell, loads at xs
          unsigned char& s = seen[(size_t)(row * nx + x)];
          if (s) return fail("cell worked twice", nx, ny, nz, cpt, item, lane);
          s = 1;
          if (x == nx - 1) owner[(size_t)row] = item;
        }
      }
    }
    for (size_t i = 0; i < seen.size(); i++)
      if (!seen[i]) return fail("cell never worked", nx, ny, nz, cpt, (long)i, -1);
    for (long row = 0; row < rows; row++)
      if (last_item_of_row(A, row) != owner[(size_t)row]) return fail("last_item_of_row", nx, ny, nz, cpt, row, -1);
    items_total += A.nwaves;
    grids++;
  }
  std::printf("ok %d %ld\n", grids, items_total);
  return 0;
}
