// cs_halo_maps.h -- index maps of the lateral ghost exchange between the six
// sectors, shared by the HIP kernels and by host code.
//
// Forward maps (pull): which owned zone / face of which sector a lateral ghost
// slot copies (the index flips of 330fetch_buffers_fill_in_ux.f:232-294 and
// its y / z counterparts, with 203fill_ghost2d_global.f:37-75's neighbours).
// Inverse maps (push): every ghost slot that copies a given owned zone / face.
// The zone and face updates call the inverse maps to write their results into
// the neighbours' ghosts as they produce them, so a stage needs no separate
// exchange; the two directions are checked against each other exhaustively.
//
// nbr[s] = sectors right, left, up, down of sector s (0-based).  The relations
// are right<->up, left<->left, down<->down: if o is the right neighbour of s
// then s is the up neighbour of o, and so on.
//
// Strips: 0 right (i > N), 1 left (i < 1), 2 up (j > N), 3 down (j < 1); depth
// a in [1, g] away from the panel edge, tt along it.  Face families: 0 x-faces
// (Bx), 1 y-faces (By), 2 z-faces (Bz).
#pragma once
#include "cs_internal.h"

// Ghost zone (strip, a, tt), tt in [1, N], of sector s: its destination
// (di, dj) and the owned zone (si, sj) of sector nbr[s][strip] it copies.
// The maps are written as selects, with no branch per case, so every lane of
// a wavefront evaluates the same straight-line code.
CS_HD void ghost_zone_map(int N, int strip, int a, int tt, int& di, int& dj, int& si, int& sj) {
  di = strip == 0 ? N + a : (strip == 1 ? 1 - a : tt);
  dj = strip == 2 ? N + a : (strip == 3 ? 1 - a : tt);
  si = strip == 0 ? tt : (strip == 1 ? a : (strip == 2 ? N + 1 - a : N + 1 - tt));
  sj = strip == 0 ? N + 1 - a : (strip == 1 ? N + 1 - tt : (strip == 2 ? tt : a));
}

// Ghost face (strip, a, fam, tt) of sector s.  fam 0 is the family normal to
// the strip (tt in [1, N], cell k), fam 1 the tangential one (tt in [0, N],
// cell k), fam 2 the z-faces (tt in [1, N], z-vertex k).  Destination family
// and index, source family and index in sector nbr[s][strip], and the sign:
//   right  fam 0  Bx(N+a, tt)  = -By(tt, N-a)      left  fam 0  Bx(-a, tt)   = -Bx(a, N+1-tt)
//          fam 1  By(N+a, tt)  = +Bx(tt, N+1-a)          fam 1  By(1-a, tt)  = -By(a, N-tt)
//          fam 2  Bz(N+a, tt)  = +Bz(tt, N+1-a)          fam 2  Bz(1-a, tt)  = +Bz(a, N+1-tt)
//   up     fam 0  By(tt, N+a)  = -Bx(N-a, tt)      down  fam 0  By(tt, -a)   = -By(N+1-tt, a)
//          fam 1  Bx(tt, N+a)  = +By(N+1-a, tt)          fam 1  Bx(tt, 1-a)  = -Bx(N-tt, a)
//          fam 2  Bz(tt, N+a)  = +Bz(N+1-a, tt)          fam 2  Bz(tt, 1-a)  = +Bz(N+1-tt, a)
CS_HD void ghost_face_map(int N, int strip, int a, int fam, int tt, int& dfam, int& di, int& dj, int& sfam,
                          int& si, int& sj, double& sign) {
  const bool f0 = fam == 0, f1 = fam == 1, f2 = fam == 2;
  const bool ew = strip < 2;  // right / left strips
  dfam = (ew || f2) ? fam : 1 - fam;
  sfam = f2 ? 2 : ((strip == 0 || strip == 3) ? 1 - fam : fam);
  sign = f2 ? 1.0 : ((strip == 0 || strip == 2) && f1 ? 1.0 : -1.0);
  di = strip == 0 ? N + a : (strip == 1 ? (f0 ? -a : 1 - a) : tt);
  dj = strip == 2 ? N + a : (strip == 3 ? (f0 ? -a : 1 - a) : tt);
  si = strip == 0 ? tt : (strip == 1 ? a : (strip == 2 ? (f0 ? N - a : N + 1 - a) : (f1 ? N - tt : N + 1 - tt)));
  sj = strip == 0 ? (f0 ? N - a : N + 1 - a) : (strip == 1 ? (f1 ? N - tt : N + 1 - tt) : (strip == 2 ? tt : a));
}

// Every lateral ghost zone that copies owned zone (i, j) of sector o:
// f(sector, di, dj).  At most two (a zone within g of two panel edges).
template <class Fn>
CS_HD void zone_pushes(const int (*nbr)[4], int N, int g, int o, int i, int j, Fn&& f) {
  if (j > N - g) f(nbr[o][2], 2 * N + 1 - j, i);  // right strip of the up neighbour
  if (i <= g) f(nbr[o][1], 1 - i, N + 1 - j);     // left strip of the left neighbour
  if (i > N - g) f(nbr[o][0], j, 2 * N + 1 - i);  // up strip of the right neighbour
  if (j <= g) f(nbr[o][3], N + 1 - i, 1 - j);     // down strip of the down neighbour
}

// The corner ghost blocks of a sector copy its own corner zones: the g x g
// ghosts that owned zone (i, j) fills, f(di, dj) (none unless (i, j) is a
// panel corner zone).
template <class Fn>
CS_HD void corner_pushes(int N, int g, int i, int j, Fn&& f) {
  if ((i != 1 && i != N) || (j != 1 && j != N)) return;
  for (int b = 1; b <= g; ++b)
    for (int a = 1; a <= g; ++a) f(i == N ? N + a : 1 - a, j == N ? N + b : 1 - b);
}

// Every lateral ghost face that copies owned face (i, j) of family fam in
// sector o: f(dest family, sector, di, dj, sign).  Owned x-faces: i in [0, N],
// j in [1, N]; y-faces: i in [1, N], j in [0, N]; z-faces: i, j in [1, N]
// (the radial index is carried unchanged).
template <class Fn>
CS_HD void face_pushes(const int (*nbr)[4], int N, int g, int fam, int o, int i, int j, Fn&& f) {
  if (fam == 0) {
    if (j > N - g) f(1, nbr[o][2], 2 * N + 1 - j, i, 1.0);
    if (i >= 1 && i <= g) f(0, nbr[o][1], -i, N + 1 - j, -1.0);
    if (i >= N - g && i <= N - 1) f(1, nbr[o][0], j, 2 * N - i, -1.0);
    if (j <= g) f(0, nbr[o][3], N - i, 1 - j, -1.0);
  } else if (fam == 1) {
    if (j >= N - g && j <= N - 1) f(0, nbr[o][2], 2 * N - j, i, -1.0);
    if (i <= g) f(1, nbr[o][1], 1 - i, N - j, -1.0);
    if (i > N - g) f(0, nbr[o][0], j, 2 * N + 1 - i, 1.0);
    if (j >= 1 && j <= g) f(1, nbr[o][3], N + 1 - i, -j, -1.0);
  } else {
    if (j > N - g) f(2, nbr[o][2], 2 * N + 1 - j, i, 1.0);
    if (i <= g) f(2, nbr[o][1], 1 - i, N + 1 - j, 1.0);
    if (i > N - g) f(2, nbr[o][0], j, 2 * N + 1 - i, 1.0);
    if (j <= g) f(2, nbr[o][3], N + 1 - i, 1 - j, 1.0);
  }
}
