// Simple update: imaginary-time evolution of an iPEPS with bond weights on a 2x2 unit cell, driven
// on the GPU tensor layer.
//
// Follows the nearest-neighbour branch of TimePropagator<Scalar, TimeScalar, HamScalar, Symmetry>
// (src/PEPS/TimePropagator.cpp): t_step (:78-105, the branch without diagonal terms), t_step_h
// (:113-190), t_step_v (:192-226), initU (:453-487, gates exp(-dt H)), SimpleUpdate::renormalize
// (src/PEPS/SimpleUpdate.cpp:5-14), applyWeights (src/PEPS/PEPSContractions.cpp:93-104) and the
// weight bookkeeping of iPEPS::initWeightTensors / updateAtensors (src/PEPS/iPEPS.cpp:575-602).
//
// Sites and bonds are indexed by x * 2 + y with x, y taken mod 2, as the CTM driver indexes its
// sites.  whs[x*2+y] is the weight on the bond between (x, y) and (x+1, y), wvs[x*2+y] the weight
// on the bond between (x, y-1) and (x, y).
#pragma once
#include <vector>

#include "tensor.hpp"

namespace xa {

// The members of iPEPS that the simple update reads and writes: the bare site tensors Gs
// (Tensor<2,3>(left, top | right, bottom, phys)), the diagonal bond weights and the bond dimension
// D that renormalize truncates to.
struct WeightedPEPS {
    std::vector<DTensor> Gs, whs, wvs;
    std::size_t D = 0;
};

// iPEPS::initWeightTensors (iPEPS.cpp:575-591): Gs = As, every weight the identity on the site's
// right (horizontal) or top (vertical) leg.
WeightedPEPS init_weight_tensors(Runtime* rt, const std::vector<DTensor>& A, std::size_t D);
// iPEPS::updateAtensors (iPEPS.cpp:593-602): As = applyWeights(Gs, sqrt of the four weights).
std::vector<DTensor> weighted_sites(const WeightedPEPS& psi);
// applyWeights (PEPSContractions.cpp:93-104): absorb wL (twisted), wT (twisted), wR and wB into the
// left, top, right and bottom legs of A.
DTensor apply_weights(const DTensor& A, const DTensor& wL, const DTensor& wT, const DTensor& wR, const DTensor& wB, int precision);

class TimePropagator {
public:
    // Hh[x*2+y], Hv[x*2+y]: the bond Hamiltonians Tensor<2,2>(p, p | p, p) of the horizontal bond
    // (x, y)-(x+1, y) and of the vertical bond (x, y)-(x, y+1) (TwoSiteObservable::data_h/data_v);
    // the gates are exp(-dt H) (initU).
    TimePropagator(Runtime* rt, const std::vector<DTensor>& Hh, const std::vector<DTensor>& Hv, double dt);

    void t_step(WeightedPEPS& psi);
    void t_step_h(WeightedPEPS& psi, int x, int y);  // bond (x, y)-(x+1, y)
    void t_step_v(WeightedPEPS& psi, int x, int y);  // bond (x, y-1)-(x, y)

    Runtime* rt = nullptr;
    double dt = 0.0;
    int precision = XH_FP64;
    std::vector<DTensor> Uh, Uv;                  // gates
    std::vector<DTensor> spectrum_h, spectrum_v;  // last un-normalised weights of each bond
    long n_bond_updates = 0;
};

}  // namespace xa
