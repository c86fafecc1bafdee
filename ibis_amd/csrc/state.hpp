// state.hpp -- conserved -> primitive conversion of one cell, shared by the
// residual's conversion kernel and the batch's fused stage kernel so that
// both evaluate the same expressions
// (FV/primative_conserved_conversion.cpp:5-51).
#ifndef IBIS_STATE_HPP
#define IBIS_STATE_HPP

#include <hip/hip_runtime.h>

namespace ibis {

// primitive rows of a flow state
enum { P_RHO = 0, P_P = 1, P_T = 2, P_E = 3, P_VX = 4, P_VY = 5, P_VZ = 6, NPRIM = 7 };

// q: rho, rho vx, rho vy, [rho vz,] rho E  ->  out[P_*]
template <class S, int DIM>
__host__ __device__ __forceinline__ void cons_to_prim_cell(const S* q, double R, double Cv,
                                                           S* out) {
    const S rho = q[0];
    const S vx = q[1] / rho;
    const S vy = q[2] / rho;
    const S vz = DIM == 3 ? q[3] / rho : S(0.0);
    const S ke = 0.5 * (vx * vx + vy * vy + vz * vz);
    const S u = q[DIM + 1] / rho - ke;
    const S temp = u / Cv;
    out[P_RHO] = rho;
    out[P_VX] = vx;
    out[P_VY] = vy;
    out[P_VZ] = vz;
    out[P_E] = u;
    out[P_T] = temp;
    out[P_P] = rho * R * temp;
}

}  // namespace ibis

#endif
