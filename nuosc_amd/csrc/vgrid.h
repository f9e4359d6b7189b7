// vgrid.h -- setup-time geometry of the nuosc core: the icosahedral Voronoi
// velocity grid (icosahedron.h:36-171) and the angular distributions of the
// initial data (nuosc_init.cpp:8-15, 74-96).  Host C++, FP64.
#pragma once
#include <vector>

namespace nuosc {

struct VGrid {
    std::vector<double> vx, vy, vz, vw;
    int size() const { return (int)vw.size(); }
};

// N = 10*nn^2 + 2 unit directions from the refined icosahedron; weights are
// the planar Voronoi areas of icosahedron.h (they do not sum to 4*pi).
VGrid make_icosahedral_vgrid(int nn);

// Same directions, equal weights 4*pi/N.
VGrid make_equal_weight_vgrid(int nn);

// Normalised angular distributions G0 = ng / sum(w ng) and
// G0b = alpha * ngb / sum(w ngb) with ng = exp(-(vx-1)^2/(2 sx^2) - (vz-1)^2/(2 sz^2))
// (nuosc_init.cpp:8-10, 74-96; normalisation summed in direction order).
void angular_distributions(const VGrid& g, double alpha, double lnue, double lnueb, double lnuex,
                           double lnuebx, std::vector<double>& G0, std::vector<double>& G0b);

}  // namespace nuosc
