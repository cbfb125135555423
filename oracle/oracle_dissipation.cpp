// =============================================================================
// oracle/oracle_dissipation.cpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
// CPU restatement of the operator-split relaxation (src2d/dissipation2d.jl).
// =============================================================================
#include "oracle_solver.hpp"

namespace orc {

void relax_dissipative(Ctx& c, double tau, double dtau) {
    (void)c; (void)tau; (void)dtau;
}

}  // namespace orc
