// Host-side check of the batched stencil's blocked site order (border_site in
// mh_bkern.h): for every lattice shape and block shape given on the command
// line, the logical index -> checkerboard index map must be a bijection of
// [0, Vh) for both parities, with the coordinates site_coords gives.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../milc_qcd_amd/csrc/mh_bkern.h"

int main(int argc, char** argv) {
  if (argc != 9) return 2;
  int dims[4], b[4];
  for (int i = 0; i < 4; i++) dims[i] = atoi(argv[1 + i]);
  for (int i = 0; i < 4; i++) b[i] = atoi(argv[5 + i]);
  mh::Geom g = mh::make_geom(dims);
  mh::BOrder o = mh::make_border(g, b[0], b[1], b[2], b[3]);
  if (!o.on) {
    printf("OFF\n");
    return 0;
  }
  for (int pb = 0; pb < 2; pb++) {
    std::vector<int> seen(g.Vh, 0);
    for (int i = 0; i < g.Vh; i++) {
      mh::Coord c;
      int cb = mh::border_site(g, o, i, pb, c);
      if (cb < 0 || cb >= g.Vh || seen[cb]) {
        printf("BAD index pb=%d i=%d cb=%d\n", pb, i, cb);
        return 1;
      }
      seen[cb] = 1;
      mh::Coord r = mh::site_coords(g, cb, pb);
      if (r.x != c.x || r.y != c.y || r.z != c.z || r.t != c.t || r.rowbase != c.rowbase) {
        printf("BAD coords pb=%d i=%d cb=%d\n", pb, i, cb);
        return 1;
      }
    }
  }
  printf("OK\n");
  return 0;
}
