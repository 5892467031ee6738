// Exhaustive check of the lateral ghost maps of cs_halo_maps.h: the inverse
// (push) maps used by the zone and face updates must write exactly the ghost
// slots the forward (pull) exchange fills, each from the same owned slot and
// with the same sign.  Built and run by tests/test_cpu_library.py.
//
//   halo_maps_check N g   ->  prints "ok <zone ghosts> <face ghosts>" or the mismatches
#include <cstdio>
#include <cstdlib>
#include <map>
#include <tuple>

#include "cs_halo_maps.h"

static const int nbr[6][4] = {{2, 5, 1, 4}, {0, 3, 2, 5}, {1, 4, 0, 3}, {4, 1, 5, 2}, {5, 2, 3, 0}, {3, 0, 4, 1}};

using Key = std::tuple<int, int, int, int>;         // family (or -1 for zones), sector, i, j
using Src = std::tuple<int, int, int, int, double>;  // family, sector, i, j, sign

static int compare(const char* what, const std::map<Key, Src>& pull, const std::map<Key, Src>& push, int dups) {
  int bad = dups;
  for (const auto& kv : pull) {
    auto it = push.find(kv.first);
    if (it == push.end() || it->second != kv.second) {
      if (bad < 10)
        printf("%s ghost (%d,%d,%d,%d): pull from (%d,%d,%d,%d,%g), push %s\n", what, std::get<0>(kv.first),
               std::get<1>(kv.first), std::get<2>(kv.first), std::get<3>(kv.first), std::get<0>(kv.second),
               std::get<1>(kv.second), std::get<2>(kv.second), std::get<3>(kv.second), std::get<4>(kv.second),
               it == push.end() ? "missing" : "differs");
      ++bad;
    }
  }
  for (const auto& kv : push)
    if (!pull.count(kv.first)) {
      if (bad < 10)
        printf("%s ghost (%d,%d,%d,%d) pushed but not pulled\n", what, std::get<0>(kv.first), std::get<1>(kv.first),
               std::get<2>(kv.first), std::get<3>(kv.first));
      ++bad;
    }
  return bad;
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const int N = atoi(argv[1]), g = atoi(argv[2]);
  // zones: strips and corner blocks
  std::map<Key, Src> zpull, zpush;
  for (int s = 0; s < 6; ++s) {
    for (int strip = 0; strip < 4; ++strip)
      for (int a = 1; a <= g; ++a)
        for (int tt = 1; tt <= N; ++tt) {
          int di, dj, si, sj;
          ghost_zone_map(N, strip, a, tt, di, dj, si, sj);
          zpull[Key(-1, s, di, dj)] = Src(-1, nbr[s][strip], si, sj, 1.0);
        }
    for (int c = 0; c < 4; ++c)
      for (int a = 1; a <= g; ++a)
        for (int b = 1; b <= g; ++b) {
          int ci = (c & 1) ? N : 1, cj = (c & 2) ? N : 1;
          zpull[Key(-1, s, (c & 1) ? N + a : 1 - a, (c & 2) ? N + b : 1 - b)] = Src(-1, s, ci, cj, 1.0);
        }
  }
  int dups = 0;
  for (int o = 0; o < 6; ++o)
    for (int j = 1; j <= N; ++j)
      for (int i = 1; i <= N; ++i) {
        zone_pushes(nbr, N, g, o, i, j, [&](int s, int di, int dj) {
          dups += zpush.count(Key(-1, s, di, dj)) ? 1 : 0;
          zpush[Key(-1, s, di, dj)] = Src(-1, o, i, j, 1.0);
        });
        corner_pushes(N, g, i, j, [&](int di, int dj) {
          dups += zpush.count(Key(-1, o, di, dj)) ? 1 : 0;
          zpush[Key(-1, o, di, dj)] = Src(-1, o, i, j, 1.0);
        });
      }
  int bad = compare("zone", zpull, zpush, dups);
  // faces
  std::map<Key, Src> fpull, fpush;
  for (int s = 0; s < 6; ++s)
    for (int strip = 0; strip < 4; ++strip)
      for (int a = 1; a <= g; ++a)
        for (int fam = 0; fam < 3; ++fam)
          for (int tt = (fam == 1 ? 0 : 1); tt <= N; ++tt) {
            int dfam, di, dj, sfam, si, sj;
            double sign;
            ghost_face_map(N, strip, a, fam, tt, dfam, di, dj, sfam, si, sj, sign);
            fpull[Key(dfam, s, di, dj)] = Src(sfam, nbr[s][strip], si, sj, sign);
          }
  dups = 0;
  for (int fam = 0; fam < 3; ++fam)
    for (int o = 0; o < 6; ++o)
      for (int j = (fam == 1 ? 0 : 1); j <= N; ++j)
        for (int i = (fam == 0 ? 0 : 1); i <= N; ++i)
          face_pushes(nbr, N, g, fam, o, i, j, [&](int gf, int s, int di, int dj, double sg) {
            dups += fpush.count(Key(gf, s, di, dj)) ? 1 : 0;
            fpush[Key(gf, s, di, dj)] = Src(fam, o, i, j, sg);
          });
  bad += compare("face", fpull, fpush, dups);
  if (bad) {
    printf("%d mismatches\n", bad);
    return 1;
  }
  printf("ok %zu %zu\n", zpull.size(), fpull.size());
  return 0;
}
