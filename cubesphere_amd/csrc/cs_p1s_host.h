// cs_p1s_host.h -- host-side tables of the separable P1 geometry, shared by
// the host evaluations (cs_p1s_host.cpp) and the P1 stage's device set-up.
#pragma once
#include <vector>

#include "../../include/cubesphere_hip.h"
#include "cs_p1geo.h"

namespace cs {

struct P1GeoHost {
  csg1::P1Geo G;  // its pointers address the vectors below
  std::vector<double> lat, rb, TX, TY, TZ;
};

// the lattice (tri-corner blocks zero), the radial vertices and the three
// face tables of the six sectors of cfg
int p1geo_host(const cs_config* cfg, P1GeoHost& H);

}  // namespace cs
