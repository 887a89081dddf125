// Host build of the forward sweeps' plane order (perm_cell in fdtdx_amd/csrc/engine_common.h) for
// the CPU tests: for ny nz zlo zhi zg given as arguments, prints "j k" of every plane position.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../fdtdx_amd/csrc/engine_common.h"

int main(int argc, char** argv) {
  if (argc != 6) return 2;
  fdtdx::StepStatic P;
  memset(&P, 0, sizeof(P));
  P.ny = atoi(argv[1]);
  P.nz = atoi(argv[2]);
  P.zlo = atoi(argv[3]);
  P.zhi = atoi(argv[4]);
  P.zg = atoi(argv[5]);
  for (int p = 0; p < P.ny * P.nz; ++p) {
    int j, k;
    fdtdx::perm_cell(P, p, j, k);
    printf("%d %d\n", j, k);
  }
  return 0;
}
