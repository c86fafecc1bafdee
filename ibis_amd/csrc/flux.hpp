// flux.hpp -- face-normal Riemann fluxes and ideal-gas closures, device side.
//
// Each function evaluates the same expression tree as the reference's
// Kokkos lambda it replaces, so that with contraction disabled the results
// are bit-identical.  The scalar type S is double for the FP64 residual and
// float for the reduced-precision flux mode.
//
// A face state is {rho, p, T, e, u, v, w} with the velocity already rotated
// into the face frame (u normal, v along t1, w along t2).  Outputs are
// {mass, normal momentum, t1 momentum, t2 momentum, energy}.
#ifndef IBIS_FLUX_HPP
#define IBIS_FLUX_HPP

#include <hip/hip_runtime.h>

#include "dual.hpp"

namespace ibis {

enum { S_RHO = 0, S_P = 1, S_T = 2, S_E = 3, S_U = 4, S_V = 5, S_W = 6, NSTATE = 7 };

struct GasConst {
    double R, Cv, Cp, gamma;
};

// Ausmdv::compute_flux, src/finite_volume/finite_volume/flux_calculators/ausmdv.cpp:9-131
template <class S, bool THREE_D>
__host__ __device__ __forceinline__ void flux_ausmdv(const S* L, const S* Rs, typename lit<S>::type R,
                                            typename lit<S>::type Cv,
                                            typename lit<S>::type gamma, S* out) {
    using K = typename lit<S>::type;
    S rL = L[S_RHO], pL = L[S_P];
    S pLrL = pL / rL;
    S uL = L[S_U], vL = L[S_V], wL = L[S_W];
    S eL = Cv * L[S_T];
    S aL = sqrt(gamma * R * L[S_T]);
    S keL = K(0.5) * (uL * uL + vL * vL + wL * wL);
    S HL = eL + pLrL + keL;
    S rR = Rs[S_RHO], pR = Rs[S_P];
    S pRrR = pR / rR;
    S uR = Rs[S_U], vR = Rs[S_V], wR = Rs[S_W];
    S eR = Cv * Rs[S_T];
    S aR = sqrt(gamma * R * Rs[S_T]);
    S keR = K(0.5) * (uR * uR + vR * vR + wR * wR);
    S HR = eR + pRrR + keR;

    S alphaL = K(2.0) * pLrL / (pLrL + pRrR);
    S alphaR = K(2.0) * pRrR / (pLrL + pRrR);
    S am = kmax(aL, aR);
    S ML = uL / am;
    S MR = uR / am;
    S pLplus, uLplus;
    S duL = K(0.5) * (uL + fabs(uL));
    if (fabs(ML) <= K(1.0)) {
        pLplus = pL * (ML + K(1.0)) * (ML + K(1.0)) * (K(2.0) - ML) * K(0.25);
        uLplus = alphaL * ((uL + am) * (uL + am) / (K(4.0) * am) - duL) + duL;
    } else {
        pLplus = pL * duL / uL;
        uLplus = duL;
    }
    S pRminus, uRminus;
    S duR = K(0.5) * (uR - fabs(uR));
    if (fabs(MR) <= K(1.0)) {
        pRminus = pR * (MR - K(1.0)) * (MR - K(1.0)) * (K(2.0) + MR) * K(0.25);
        uRminus = alphaR * (-(uR - am) * (uR - am) / (K(4.0) * am) - duR) + duR;
    } else {
        pRminus = pR * duR / uR;
        uRminus = duR;
    }
    S ru_half = uLplus * rL + uRminus * rR;
    S p_half = pLplus + pRminus;
    S dp = pL - pR;
    dp = K(10.0) * fabs(dp) / kmin(pL, pR);
    S s = K(0.5) * kmin(K(1.0), dp);
    S ru2_AUSMV = uLplus * rL * uL + uRminus * rR * uR;
    S ru2_AUSMD = K(0.5) * (ru_half * (uL + uR) - fabs(ru_half) * (uR - uL));
    S ru2_half = (K(0.5) + s) * ru2_AUSMV + (K(0.5) - s) * ru2_AUSMD;
    S m = ru_half, mx = ru2_half + p_half, my, mz = K(0.0), en;
    if (ru_half >= K(0.0)) {
        my = ru_half * vL;
        if (THREE_D) mz = ru_half * wL;
        en = ru_half * HL;
    } else {
        my = ru_half * vR;
        if (THREE_D) mz = ru_half * wR;
        en = ru_half * HR;
    }
    const K C_EFIX = K(0.125);
    bool caseA = ((uL - aL) < K(0.0)) && ((uR - aR) > K(0.0));
    bool caseB = ((uL + aL) < K(0.0)) && ((uR + aR) > K(0.0));
    S d_ua = K(0.0);
    if (caseA && !caseB) d_ua = C_EFIX * ((uR - aR) - (uL - aL));
    if (caseB && !caseA) d_ua = C_EFIX * ((uR + aR) - (uL + aL));
    if (d_ua != K(0.0)) {
        m -= d_ua * (rR - rL);
        mx -= d_ua * (rR * uR - rL * uL);
        my -= d_ua * (rR * vR - rL * vL);
        if (THREE_D) mz -= d_ua * (rR * wR - rL * wL);
        en -= d_ua * (rR * HR - rL * HL);
    }
    out[0] = m;
    out[1] = mx;
    out[2] = my;
    out[3] = mz;
    out[4] = en;
}

// Hanel::compute_flux, flux_calculators/hanel.cpp:8-64
template <class S, bool THREE_D>
__host__ __device__ __forceinline__ void flux_hanel(const S* L, const S* Rs, typename lit<S>::type R,
                                           typename lit<S>::type Cv,
                                           typename lit<S>::type gamma, S* out) {
    using K = typename lit<S>::type;
    S rL = L[S_RHO], pL = L[S_P];
    S eL = Cv * L[S_T];
    S aL = sqrt(gamma * R * L[S_T]);
    S uL = L[S_U], vL = L[S_V], wL = L[S_W];
    S keL = K(0.5) * (uL * uL + vL * vL + wL * wL);
    S pLrL = pL / rL;
    S HL = eL + pLrL + keL;
    S rR = Rs[S_RHO], pR = Rs[S_P];
    S eR = Cv * Rs[S_T];
    S aR = sqrt(gamma * R * Rs[S_T]);
    S uR = Rs[S_U], vR = Rs[S_V], wR = Rs[S_W];
    S keR = K(0.5) * (uR * uR + vR * vR + wR * wR);
    S pRrR = pR / rR;
    S HR = eR + pRrR + keR;
    S pLplus, uLplus;
    if (fabs(uL) <= aL) {
        uLplus = K(1.0) / (K(4.0) * aL) * (uL + aL) * (uL + aL);
        pLplus = pL * uLplus * (K(1.0) / aL * (K(2.0) - uL / aL));
    } else {
        uLplus = K(0.5) * (uL + fabs(uL));
        pLplus = pL * uLplus * (K(1.0) / uL);
    }
    S pRminus, uRminus;
    if (fabs(uR) <= aR) {
        uRminus = K(-1.0) / (K(4.0) * aR) * (uR - aR) * (uR - aR);
        pRminus = pR * uRminus * (K(1.0) / aR * (K(-2.0) - uR / aR));
    } else {
        uRminus = K(0.5) * (uR - fabs(uR));
        pRminus = pR * uRminus * (K(1.0) / uR);
    }
    S p_half = pLplus + pRminus;
    out[0] = uLplus * rL + uRminus * rR;
    out[1] = uLplus * rL * uL + uRminus * rR * uR + p_half;
    out[2] = uLplus * rL * vL + uRminus * rR * vR;
    out[3] = THREE_D ? (uLplus * rL * wL + uRminus * rR * wR) : K(0.0);
    out[4] = uLplus * rL * HL + uRminus * rR * HR;
}

// Ldfss2::compute_flux, flux_calculators/ldfss2.cpp:10-89
template <class S, bool THREE_D>
__host__ __device__ __forceinline__ void flux_ldfss2(const S* L, const S* Rs, typename lit<S>::type R,
                                            typename lit<S>::type Cv,
                                            typename lit<S>::type gamma,
                                            typename lit<S>::type delta, S* out) {
    using K = typename lit<S>::type;
    S rL = L[S_RHO], pL = L[S_P];
    S pLrL = pL / rL;
    S uL = L[S_U], vL = L[S_V], wL = L[S_W];
    S eL = Cv * L[S_T];
    S aL = sqrt(gamma * R * L[S_T]);
    S keL = K(0.5) * (uL * uL + vL * vL + wL * wL);
    S HL = eL + pLrL + keL;
    S rR = Rs[S_RHO], pR = Rs[S_P];
    S pRrR = pR / rR;
    S uR = Rs[S_U], vR = Rs[S_V], wR = Rs[S_W];
    S eR = Cv * Rs[S_T];
    S aR = sqrt(gamma * R * Rs[S_T]);
    S keR = K(0.5) * (uR * uR + vR * vR + wR * wR);
    S HR = eR + pRrR + keR;
    S am = K(0.5) * (aL + aR);
    S ML = uL / am;
    S MR = uR / am;
    S MpL = K(0.25) * (ML + K(1.0)) * (ML + K(1.0));
    S MmR = K(-0.25) * (MR - K(1.0)) * (MR - K(1.0));
    S alphaL = K(0.5) * (K(1.0) + copysign(S(1.0), ML));
    S alphaR = K(0.5) * (K(1.0) - copysign(S(1.0), MR));
    // S(...): on the host side float floor/fabs may widen to double
    S betaL = -kmax(K(0.0), S(K(1.0) - S(floor(fabs(ML)))));
    S betaR = -kmax(K(0.0), S(K(1.0) - S(floor(fabs(MR)))));
    S PL = K(0.25) * (ML + K(1.0)) * (ML + K(1.0)) * (K(2.0) - ML);
    S PR = K(0.25) * (MR - K(1.0)) * (MR - K(1.0)) * (K(2.0) + MR);
    S DL = alphaL * (K(1.0) + betaL) - betaL * PL;
    S DR = alphaR * (K(1.0) + betaR) - betaR * PR;
    S Mhalf = K(0.25) * betaL * betaR * pow(sqrt(K(0.5) * (ML * ML + MR * MR)) - K(1.0), S(2.0));
    S MhalfL = Mhalf * (K(1.0) - ((pL - pR) / (pL + pR) + delta * (fabs(pL - pR) / pL)));
    S MhalfR = Mhalf * (K(1.0) + ((pL - pR) / (pL + pR) - delta * (fabs(pL - pR) / pR)));
    S CL = alphaL * (K(1.0) + betaL) * ML - betaL * MpL - MhalfL;
    S CR = alphaR * (K(1.0) + betaR) * MR - betaR * MmR + MhalfR;
    S ru_half = am * rL * CL + am * rR * CR;
    S ru2_half = am * rL * CL * uL + am * rR * CR * uR;
    S p_half = DL * pL + DR * pR;
    out[0] = ru_half;
    out[1] = ru2_half + p_half;
    out[2] = am * rL * CL * vL + am * rR * CR * vR;
    out[3] = THREE_D ? (am * rL * CL * wL + am * rR * CR * wR) : K(0.0);
    out[4] = am * rL * CL * HL + am * rR * CR * HR;
}

// Rusanov::compute_flux, flux_calculators/rusanov.cpp:10-52
template <class S, bool THREE_D>
__host__ __device__ __forceinline__ void flux_rusanov(const S* L, const S* Rs, typename lit<S>::type R,
                                             typename lit<S>::type Cv,
                                             typename lit<S>::type gamma, S* out) {
    using K = typename lit<S>::type;
    S rL = L[S_RHO];
    S eL = Cv * L[S_T];
    S pL = L[S_P];
    S uL = L[S_U], vL = L[S_V], wL = L[S_W];
    S ruL = rL * uL;
    S keL = K(0.5) * (uL * uL + vL * vL + wL * wL);
    S HL = eL + pL / rL + keL;
    S aL = sqrt(gamma * R * L[S_T]);
    S rR = Rs[S_RHO];
    S eR = Cv * Rs[S_T];
    S pR = Rs[S_P];
    S uR = Rs[S_U], vR = Rs[S_V], wR = Rs[S_W];
    S ruR = rR * uR;
    S keR = K(0.5) * (uR * uR + vR * vR + wR * wR);
    S HR = eR + pR / rR + keR;
    S aR = sqrt(gamma * R * Rs[S_T]);
    S S_plus = kmax(S(fabs(uL - aL)), S(fabs(uR - aR)));
    S_plus = kmax(S_plus, S(fabs(uL + aL)));
    S_plus = kmax(S_plus, S(fabs(uR + aR)));
    out[0] = K(0.5) * (ruL + ruR) - K(0.5) * S_plus * (rR - rL);
    out[1] = K(0.5) * (ruL * uL + pL + ruR * uR + pR) - K(0.5) * S_plus * (rR * uR - rL * uL);
    out[2] = K(0.5) * (ruL * vL + ruR * vR) - K(0.5) * S_plus * (rR * vR - rL * vL);
    out[3] = THREE_D ? (K(0.5) * (ruL * wL + ruR * wR) - K(0.5) * S_plus * (rR * wR - rL * wL))
                     : K(0.0);
    out[4] = K(0.5) * (ruL * HL + ruR * HR) - K(0.5) * S_plus * (rR * HR - rL * HL);
}

// Thermodynamic closure after reconstruction (src/gas/gas/gas_model.h:91-116)
template <class S>
__host__ __device__ __forceinline__ void thermo_update(int mode, typename lit<S>::type R,
                                              typename lit<S>::type Cv, S& rho, S& p, S& T,
                                              S& e) {
    switch (mode) {
        case 0:  // rho_p: update_thermo_from_rhop
            T = p / (rho * R);
            e = Cv * T;
            break;
        case 1:  // rho_T: update_thermo_from_rhoT
            p = rho * R * T;
            e = Cv * T;
            break;
        case 2:  // rho_u: update_thermo_from_rhou
            T = e / Cv;
            p = rho * R * T;
            break;
        default:  // p_T: update_thermo_from_pT
            rho = p / (R * T);
            e = Cv * T;
            break;
    }
}

}  // namespace ibis
#endif
