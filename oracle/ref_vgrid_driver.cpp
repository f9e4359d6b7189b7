// Test-infrastructure driver (not product code): links the reference's own
// self-contained velocity-grid header (icosahedron.h + Point.h, compiled where
// they lie) and prints the grid as exact hex floats, one direction per line:
//   vx vy vz vw
// Usage: ref_vgrid <nn>   (N = 10*nn*nn + 2 directions)
#include <cstdio>
#include <cstdlib>
#include "icosahedron.h"

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: %s nn\n", argv[0]); return 2; }
    int nn = std::atoi(argv[1]);
    if (nn < 1 || nn > 64) { std::fprintf(stderr, "nn out of range\n"); return 2; }
    IcosahedronVoronoi ico(nn);
    for (unsigned k = 0; k < ico.N; ++k)
        std::printf("%a %a %a %a\n", ico.X[k].x, ico.X[k].y, ico.X[k].z, ico.vw[k]);
    return 0;
}
