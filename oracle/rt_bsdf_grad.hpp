// ORACLE — test infrastructure only (see rt_core.hpp).
// d f / d parameter of a material's BSDF at fixed directions, for the oracle's adjoint (no
// counterpart in the reference): the BSDF of rt_bsdf.cpp (after bsdf.rs:384-996 and
// bsdf/energy.rs) rebuilt and evaluated with one parameter (roughness, base colour r, g, b or IOR)
// carried as a dual number (value, derivative), so the derivative is exact for the branch and the
// energy-table cell the value is in. Pinned against Richardson-extrapolated central differences
// of Bsdf::surface / Bsdf::eval away from table nodes and total internal reflection
// (tests/test_oracle_adjoint.py).
#pragma once
#include "rt_bsdf.hpp"

namespace orc {

// out[k] = d f / d parameter k (0 roughness, 1..3 base colour r, g, b, 4 IOR) at (wo, wi).
void bsdf_param_derivs_exact(const Material& m, bool entering, Vec3 wo, Vec3 wi, Lobes lobes, Rgb out[5]);

}  // namespace orc
