// Shape-sensitivity tables of the cut elements with respect to the radius of
// the inclusion that cuts them (see pgf_sens.cpp).
#pragma once

#include <cstdint>
#include <vector>

namespace pgf {

class Mesh;

struct ShapeSensTables {
    std::vector<double> vel;     // [n_enrich][3] velocity of each enrichment node
    std::vector<double> dgrad;   // layout of Mesh::vpt_grad: d(gradient)/dr of every slot
    std::vector<double> divv;    // [n_vpt] divergence of the velocity at each point
    std::vector<int32_t> param;  // [n_cut] inclusion whose radius moves the element
};

// as_reference: the parent-slot gradient derivative as the reference computes
// it (src/elem.cpp:186-221) instead of the exact material derivative
void build_shape_sensitivity(const Mesh& mesh, bool as_reference, ShapeSensTables& out);

}  // namespace pgf
