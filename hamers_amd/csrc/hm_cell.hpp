// hm_cell.hpp -- primitive state and sound speed of one cell from its
// conservative variables, single-species model: the arithmetic of the
// derived-data kernel, shared with the stable-dt kernel so that both see the
// same numbers bit for bit.
#pragma once

#include "hm_numerics.hpp"

namespace hm {

// q(e, at): conservative variable e of the cell at index `at`.
// FlowModelSingleSpecies.cpp:2719 (u), :2934 (internal energy),
// EquationOfStateIdealGas.cpp:5477 (p), :5806 (c).
template <int D, class QF>
__device__ __forceinline__ void cell_state_single(QF&& q, uint32_t iq, double gam, double& rho, double (&u)[3],
                                                  double& p, double& c) {
    rho = q(0, iq);
    const double rr = rcp_rn(rho);  // quot(a, rho, rr) rounds exactly as a / rho
    u[0] = u[1] = u[2] = 0.0;
#pragma unroll
    for (int b = 0; b < D; b++) u[b] = quot(q(1 + b, iq), rho, rr);
    double ke;
    if constexpr (D == 1)
        ke = u[0] * u[0];
    else if constexpr (D == 2)
        ke = u[0] * u[0] + u[1] * u[1];
    else
        ke = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double E = q(D + 1, iq);
    const double eps_int = quot(E, rho, rr) - double(1) / double(2) * ke;
    p = (gam - double(1)) * rho * eps_int;
    c = sqrt(quot(gam * p, rho, rr));
}

}  // namespace hm
