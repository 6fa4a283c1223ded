// Fused build_mass on trilinear hex geometry: qdata straight from the vertex
// coordinates, without the gradient Q-vector of the unfused composition
// (9 doubles per quadrature point: 1.6 GB for 40^3 hexes at Q=7).
#pragma once

#include <hip/hip_runtime.h>

#include "basis.hpp"
#include "restriction.hpp"

namespace femhip {

// Whether fused_build_mass handles this mesh restriction and basis: an offset
// restriction of 8-node elements with 3 components, and a dim-3, 3-component,
// P=2 basis.
bool fused_build_mass_supported(const Restriction& mesh_rstr, const Basis& mesh_basis);

// qdata[q*E + e] = 0 + det(J) * w_q for every quadrature point q of every
// element e, bit for bit what operator_apply computes for the build_mass
// operator (include/qfunctions/mass.hpp:18-40) with inputs dx (Grad of the
// coordinates through mesh_rstr and mesh_basis) and weights (Weight of
// mesh_basis), and output qdata on a backend-strided restriction
// (src/operator.cpp:148-182): the same contractions in the same order, the
// same cofactor expansion and the same weight product, none with fused
// multiply-add.
void fused_build_mass(const Restriction& mesh_rstr, const Basis& mesh_basis, const double* coords, double* qdata,
                      hipStream_t s);

}  // namespace femhip
