// flux_adj.hpp -- reverse-mode (vector-Jacobian) products of the flux
// calculators and of the thermo closure, written out by hand.
//
// Each *_vjp function re-evaluates the forward flux of flux.hpp in double
// (recomputation instead of storage) and then accumulates
//   Lb += (dF/dL)^T out_b,  Rb += (dF/dR)^T out_b
// for the face-frame states L, R = (rho, p, T, e, u, v, w) (e does not
// enter the fluxes: they use Cv * T).  Every branch (subsonic splitting,
// entropy fix, max/min, |x|) is taken as the reference's Dual arithmetic
// takes it (src/util/util/dual.h:45-73, 214-296): comparisons on the value,
// max/min returning the second argument on ties, d|x|/dx = -1 for x < 0 and
// +1 otherwise, so that a dot product against the Dual tangent residual is
// exact up to rounding.
#ifndef IBIS_FLUX_ADJ_HPP
#define IBIS_FLUX_ADJ_HPP

#include <hip/hip_runtime.h>

#include "flux.hpp"

namespace ibis {

__host__ __device__ __forceinline__ double sgn1(double x) { return x < 0.0 ? -1.0 : 1.0; }

// Ausmdv::compute_flux reverse (ausmdv.cpp:9-131)
template <bool THREE_D>
__host__ __device__ inline void flux_ausmdv_vjp(const double* L, const double* Rs, double R,
                                                double Cv, double gamma, const double* ob,
                                                double* Lb, double* Rb) {
    const double rL = L[S_RHO], pL = L[S_P], TL = L[S_T], uL = L[S_U], vL = L[S_V], wL = L[S_W];
    const double rR = Rs[S_RHO], pR = Rs[S_P], TR = Rs[S_T], uR = Rs[S_U], vR = Rs[S_V],
                 wR = Rs[S_W];
    // forward
    const double pLrL = pL / rL, eL = Cv * TL, aL = sqrt(gamma * R * TL);
    const double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
    const double HL = eL + pLrL + keL;
    const double pRrR = pR / rR, eR = Cv * TR, aR = sqrt(gamma * R * TR);
    const double keR = 0.5 * (uR * uR + vR * vR + wR * wR);
    const double HR = eR + pRrR + keR;
    const double sum = pLrL + pRrR;
    const double alphaL = 2.0 * pLrL / sum, alphaR = 2.0 * pRrR / sum;
    const bool am_is_L = aL > aR;  // Dual max: second argument on ties
    const double am = am_is_L ? aL : aR;
    const double ML = uL / am, MR = uR / am;
    const double duL = 0.5 * (uL + fabs(uL));
    const bool subL = fabs(ML) <= 1.0;
    const double YL = (uL + am) * (uL + am) / (4.0 * am);
    // the split pressures themselves are not needed: their cotangents
    // are formed from the derivative expressions below
    const double uLplus = subL ? alphaL * (YL - duL) + duL : duL;
    const double duR = 0.5 * (uR - fabs(uR));
    const bool subR = fabs(MR) <= 1.0;
    const double XR = -(uR - am) * (uR - am) / (4.0 * am);
    const double uRminus = subR ? alphaR * (XR - duR) + duR : duR;
    const double ru_half = uLplus * rL + uRminus * rR;
    const double dp = pL - pR;
    const bool pmin_is_L = pL < pR;  // Dual min: second argument on ties
    const double pmin = pmin_is_L ? pL : pR;
    const double dp2 = 10.0 * fabs(dp) / pmin;
    const bool s_const = 1.0 < dp2;
    const double s = 0.5 * (s_const ? 1.0 : dp2);
    const double ausmv = uLplus * rL * uL + uRminus * rR * uR;
    const double ausmd = 0.5 * (ru_half * (uL + uR) - fabs(ru_half) * (uR - uL));
    const bool pos = ru_half >= 0.0;
    const double C = 0.125;
    const bool caseA = ((uL - aL) < 0.0) && ((uR - aR) > 0.0);
    const bool caseB = ((uL + aL) < 0.0) && ((uR + aR) > 0.0);
    double d_ua = 0.0;
    int fix = 0;
    if (caseA && !caseB) {
        d_ua = C * ((uR - aR) - (uL - aL));
        fix = 1;
    }
    if (caseB && !caseA) {
        d_ua = C * ((uR + aR) - (uL + aL));
        fix = 2;
    }
    // reverse
    const double mb = ob[0], mxb = ob[1], myb = ob[2], mzb = THREE_D ? ob[3] : 0.0, enb = ob[4];
    double rLb = 0, pLb = 0, TLb = 0, uLb = 0, vLb = 0, wLb = 0;
    double rRb = 0, pRb = 0, TRb = 0, uRb = 0, vRb = 0, wRb = 0;
    double HLb = 0, HRb = 0, aLb = 0, aRb = 0;
    if (d_ua != 0.0) {
        const double db = -mb * (rR - rL) - mxb * (rR * uR - rL * uL) - myb * (rR * vR - rL * vL) -
                          mzb * (rR * wR - rL * wL) - enb * (rR * HR - rL * HL);
        rRb -= d_ua * (mb + mxb * uR + myb * vR + mzb * wR + enb * HR);
        rLb += d_ua * (mb + mxb * uL + myb * vL + mzb * wL + enb * HL);
        uRb -= d_ua * mxb * rR;
        uLb += d_ua * mxb * rL;
        vRb -= d_ua * myb * rR;
        vLb += d_ua * myb * rL;
        wRb -= d_ua * mzb * rR;
        wLb += d_ua * mzb * rL;
        HRb -= d_ua * enb * rR;
        HLb += d_ua * enb * rL;
        if (fix == 1) {
            uRb += C * db;
            aRb -= C * db;
            uLb -= C * db;
            aLb += C * db;
        } else {
            uRb += C * db;
            aRb += C * db;
            uLb -= C * db;
            aLb -= C * db;
        }
    }
    double rhb = mb;  // ru_half
    if (pos) {
        rhb += myb * vL + mzb * wL + enb * HL;
        vLb += myb * ru_half;
        wLb += mzb * ru_half;
        HLb += enb * ru_half;
    } else {
        rhb += myb * vR + mzb * wR + enb * HR;
        vRb += myb * ru_half;
        wRb += mzb * ru_half;
        HRb += enb * ru_half;
    }
    const double r2b = mxb;  // ru2_half
    const double phb = mxb;  // p_half
    const double sb = r2b * (ausmv - ausmd);
    const double ausmvb = r2b * (0.5 + s);
    const double ausmdb = r2b * (0.5 - s);
    rhb += 0.5 * ausmdb * ((uL + uR) - sgn1(ru_half) * (uR - uL));
    uLb += 0.5 * ausmdb * (ru_half + fabs(ru_half));
    uRb += 0.5 * ausmdb * (ru_half - fabs(ru_half));
    double uLpb = ausmvb * rL * uL, uRmb = ausmvb * rR * uR;
    rLb += ausmvb * uLplus * uL;
    uLb += ausmvb * uLplus * rL;
    rRb += ausmvb * uRminus * uR;
    uRb += ausmvb * uRminus * rR;
    if (!s_const) {
        const double dp2b = 0.5 * sb;
        const double dpb = dp2b * 10.0 / pmin * sgn1(dp);
        const double pminb = -dp2b * dp2 / pmin;
        pLb += dpb;
        pRb -= dpb;
        if (pmin_is_L)
            pLb += pminb;
        else
            pRb += pminb;
    }
    const double pLpb = phb, pRmb = phb;
    uLpb += rhb * rL;
    rLb += rhb * uLplus;
    uRmb += rhb * rR;
    rRb += rhb * uRminus;
    double amb = 0.0, alphaLb = 0.0, alphaRb = 0.0, MLb = 0.0, MRb = 0.0, duLb = 0.0, duRb = 0.0;
    // right splitting
    if (subR) {
        pRb += pRmb * (MR - 1.0) * (MR - 1.0) * (2.0 + MR) * 0.25;
        MRb += pRmb * pR * 0.25 * (2.0 * (MR - 1.0) * (2.0 + MR) + (MR - 1.0) * (MR - 1.0));
        alphaRb += uRmb * (XR - duR);
        const double XRb = uRmb * alphaR;
        duRb += uRmb * (1.0 - alphaR);
        uRb += XRb * (-2.0 * (uR - am) / (4.0 * am));
        amb += XRb * (2.0 * (uR - am) / (4.0 * am) - XR / am);
    } else {
        pRb += pRmb * duR / uR;
        duRb += pRmb * pR / uR;
        uRb += -pRmb * pR * duR / (uR * uR);
        duRb += uRmb;
    }
    uRb += 0.5 * duRb * (1.0 - sgn1(uR));
    // left splitting
    if (subL) {
        pLb += pLpb * (ML + 1.0) * (ML + 1.0) * (2.0 - ML) * 0.25;
        MLb += pLpb * pL * 0.25 * (2.0 * (ML + 1.0) * (2.0 - ML) - (ML + 1.0) * (ML + 1.0));
        alphaLb += uLpb * (YL - duL);
        const double YLb = uLpb * alphaL;
        duLb += uLpb * (1.0 - alphaL);
        uLb += YLb * (2.0 * (uL + am) / (4.0 * am));
        amb += YLb * (2.0 * (uL + am) / (4.0 * am) - YL / am);
    } else {
        pLb += pLpb * duL / uL;
        duLb += pLpb * pL / uL;
        uLb += -pLpb * pL * duL / (uL * uL);
        duLb += uLpb;
    }
    uLb += 0.5 * duLb * (1.0 + sgn1(uL));
    uLb += MLb / am;
    uRb += MRb / am;
    amb -= MLb * ML / am + MRb * MR / am;
    if (am_is_L)
        aLb += amb;
    else
        aRb += amb;
    const double sumb = -(alphaLb * alphaL + alphaRb * alphaR) / sum;
    double pLrLb = 2.0 * alphaLb / sum + sumb + HLb;
    double pRrRb = 2.0 * alphaRb / sum + sumb + HRb;
    // H = e + p/rho + ke, e = Cv T, a = sqrt(gamma R T)
    TLb += Cv * HLb + aLb * gamma * R / (2.0 * aL);
    TRb += Cv * HRb + aRb * gamma * R / (2.0 * aR);
    uLb += HLb * uL;
    vLb += HLb * vL;
    wLb += HLb * wL;
    uRb += HRb * uR;
    vRb += HRb * vR;
    wRb += HRb * wR;
    pLb += pLrLb / rL;
    rLb -= pLrLb * pLrL / rL;
    pRb += pRrRb / rR;
    rRb -= pRrRb * pRrR / rR;
    Lb[S_RHO] += rLb;
    Lb[S_P] += pLb;
    Lb[S_T] += TLb;
    Lb[S_U] += uLb;
    Lb[S_V] += vLb;
    Lb[S_W] += wLb;
    Rb[S_RHO] += rRb;
    Rb[S_P] += pRb;
    Rb[S_T] += TRb;
    Rb[S_U] += uRb;
    Rb[S_V] += vRb;
    Rb[S_W] += wRb;
}

// Hanel::compute_flux reverse (hanel.cpp:8-64)
template <bool THREE_D>
__host__ __device__ inline void flux_hanel_vjp(const double* L, const double* Rs, double R,
                                               double Cv, double gamma, const double* ob,
                                               double* Lb, double* Rb) {
    const double rL = L[S_RHO], pL = L[S_P], TL = L[S_T], uL = L[S_U], vL = L[S_V], wL = L[S_W];
    const double rR = Rs[S_RHO], pR = Rs[S_P], TR = Rs[S_T], uR = Rs[S_U], vR = Rs[S_V],
                 wR = Rs[S_W];
    const double aL = sqrt(gamma * R * TL), aR = sqrt(gamma * R * TR);
    const double pLrL = pL / rL, pRrR = pR / rR;
    const double HL = Cv * TL + pLrL + 0.5 * (uL * uL + vL * vL + wL * wL);
    const double HR = Cv * TR + pRrR + 0.5 * (uR * uR + vR * vR + wR * wR);
    const bool subL = fabs(uL) <= aL, subR = fabs(uR) <= aR;
    double uLplus, DL = 0.0, uRminus, ER = 0.0;
    if (subL) {
        uLplus = 1.0 / (4.0 * aL) * (uL + aL) * (uL + aL);
        DL = 1.0 / aL * (2.0 - uL / aL);
    } else {
        uLplus = 0.5 * (uL + fabs(uL));
    }
    if (subR) {
        uRminus = -1.0 / (4.0 * aR) * (uR - aR) * (uR - aR);
        ER = 1.0 / aR * (-2.0 - uR / aR);
    } else {
        uRminus = 0.5 * (uR - fabs(uR));
    }
    const double PL = uLplus * rL, QR = uRminus * rR;
    const double o0 = ob[0], o1 = ob[1], o2 = ob[2], o3 = THREE_D ? ob[3] : 0.0, o4 = ob[4];
    double PLb = o0 + o1 * uL + o2 * vL + o3 * wL + o4 * HL;
    double QRb = o0 + o1 * uR + o2 * vR + o3 * wR + o4 * HR;
    double uLb = o1 * PL, vLb = o2 * PL, wLb = o3 * PL, HLb = o4 * PL;
    double uRb = o1 * QR, vRb = o2 * QR, wRb = o3 * QR, HRb = o4 * QR;
    const double phb = o1;
    double uLpb = PLb * rL, rLb = PLb * uLplus;
    double uRmb = QRb * rR, rRb = QRb * uRminus;
    double pLb = 0.0, pRb = 0.0, aLb = 0.0, aRb = 0.0;
    if (subL) {
        pLb += phb * uLplus * DL;
        uLpb += phb * pL * DL;
        const double DLb = phb * pL * uLplus;
        uLb += DLb * (-1.0 / (aL * aL));
        aLb += DLb * (-2.0 / (aL * aL) + 2.0 * uL / (aL * aL * aL));
        const double c = 1.0 / (4.0 * aL);
        uLb += uLpb * 2.0 * c * (uL + aL);
        aLb += uLpb * (2.0 * c * (uL + aL) - uLplus / aL);
    } else {
        pLb += phb * uLplus / uL;
        uLpb += phb * pL / uL;
        uLb += -phb * pL * uLplus / (uL * uL);
        uLb += uLpb * 0.5 * (1.0 + sgn1(uL));
    }
    if (subR) {
        pRb += phb * uRminus * ER;
        uRmb += phb * pR * ER;
        const double ERb = phb * pR * uRminus;
        uRb += ERb * (-1.0 / (aR * aR));
        aRb += ERb * (2.0 / (aR * aR) + 2.0 * uR / (aR * aR * aR));
        const double c = -1.0 / (4.0 * aR);
        uRb += uRmb * 2.0 * c * (uR - aR);
        aRb += uRmb * (-2.0 * c * (uR - aR) - uRminus / aR);
    } else {
        pRb += phb * uRminus / uR;
        uRmb += phb * pR / uR;
        uRb += -phb * pR * uRminus / (uR * uR);
        uRb += uRmb * 0.5 * (1.0 - sgn1(uR));
    }
    double TLb = Cv * HLb + aLb * gamma * R / (2.0 * aL);
    double TRb = Cv * HRb + aRb * gamma * R / (2.0 * aR);
    uLb += HLb * uL;
    vLb += HLb * vL;
    wLb += HLb * wL;
    uRb += HRb * uR;
    vRb += HRb * vR;
    wRb += HRb * wR;
    pLb += HLb / rL;
    rLb -= HLb * pLrL / rL;
    pRb += HRb / rR;
    rRb -= HRb * pRrR / rR;
    Lb[S_RHO] += rLb;
    Lb[S_P] += pLb;
    Lb[S_T] += TLb;
    Lb[S_U] += uLb;
    Lb[S_V] += vLb;
    Lb[S_W] += wLb;
    Rb[S_RHO] += rRb;
    Rb[S_P] += pRb;
    Rb[S_T] += TRb;
    Rb[S_U] += uRb;
    Rb[S_V] += vRb;
    Rb[S_W] += wRb;
}

// Rusanov::compute_flux reverse (rusanov.cpp:10-52)
template <bool THREE_D>
__host__ __device__ inline void flux_rusanov_vjp(const double* L, const double* Rs, double R,
                                                 double Cv, double gamma, const double* ob,
                                                 double* Lb, double* Rb) {
    const double rL = L[S_RHO], pL = L[S_P], TL = L[S_T], uL = L[S_U], vL = L[S_V], wL = L[S_W];
    const double rR = Rs[S_RHO], pR = Rs[S_P], TR = Rs[S_T], uR = Rs[S_U], vR = Rs[S_V],
                 wR = Rs[S_W];
    const double ruL = rL * uL, ruR = rR * uR;
    const double HL = Cv * TL + pL / rL + 0.5 * (uL * uL + vL * vL + wL * wL);
    const double HR = Cv * TR + pR / rR + 0.5 * (uR * uR + vR * vR + wR * wR);
    const double aL = sqrt(gamma * R * TL), aR = sqrt(gamma * R * TR);
    // S+ = max(|uL-aL|, |uR-aR|, |uL+aL|, |uR+aR|) in that order (Dual max)
    const double c[4] = {uL - aL, uR - aR, uL + aL, uR + aR};
    int arg = 0;
    double S = fabs(c[0]);
    for (int k = 1; k < 4; k++)
        if (!(S > fabs(c[k]))) {
            S = fabs(c[k]);
            arg = k;
        }
    const double o0 = ob[0], o1 = ob[1], o2 = ob[2], o3 = THREE_D ? ob[3] : 0.0, o4 = ob[4];
    const double Sb = -0.5 * (o0 * (rR - rL) + o1 * (rR * uR - rL * uL) + o2 * (rR * vR - rL * vL) +
                              o3 * (rR * wR - rL * wL) + o4 * (rR * HR - rL * HL));
    double ruLb = 0.5 * (o0 + o1 * uL + o2 * vL + o3 * wL + o4 * HL);
    double ruRb = 0.5 * (o0 + o1 * uR + o2 * vR + o3 * wR + o4 * HR);
    double rLb = 0.5 * S * (o0 + o1 * uL + o2 * vL + o3 * wL + o4 * HL);
    double rRb = -0.5 * S * (o0 + o1 * uR + o2 * vR + o3 * wR + o4 * HR);
    double uLb = 0.5 * o1 * ruL + 0.5 * S * o1 * rL;
    double uRb = 0.5 * o1 * ruR - 0.5 * S * o1 * rR;
    double vLb = 0.5 * o2 * ruL + 0.5 * S * o2 * rL;
    double vRb = 0.5 * o2 * ruR - 0.5 * S * o2 * rR;
    double wLb = 0.5 * o3 * ruL + 0.5 * S * o3 * rL;
    double wRb = 0.5 * o3 * ruR - 0.5 * S * o3 * rR;
    double HLb = 0.5 * o4 * ruL + 0.5 * S * o4 * rL;
    double HRb = 0.5 * o4 * ruR - 0.5 * S * o4 * rR;
    double pLb = 0.5 * o1, pRb = 0.5 * o1;
    rLb += ruLb * uL;
    uLb += ruLb * rL;
    rRb += ruRb * uR;
    uRb += ruRb * rR;
    const double cb = Sb * sgn1(c[arg]);
    double aLb = 0.0, aRb = 0.0;
    if (arg == 0) {
        uLb += cb;
        aLb -= cb;
    } else if (arg == 1) {
        uRb += cb;
        aRb -= cb;
    } else if (arg == 2) {
        uLb += cb;
        aLb += cb;
    } else {
        uRb += cb;
        aRb += cb;
    }
    double TLb = Cv * HLb + aLb * gamma * R / (2.0 * aL);
    double TRb = Cv * HRb + aRb * gamma * R / (2.0 * aR);
    uLb += HLb * uL;
    vLb += HLb * vL;
    wLb += HLb * wL;
    uRb += HRb * uR;
    vRb += HRb * vR;
    wRb += HRb * wR;
    pLb += HLb / rL;
    rLb -= HLb * pL / (rL * rL);
    pRb += HRb / rR;
    rRb -= HRb * pR / (rR * rR);
    Lb[S_RHO] += rLb;
    Lb[S_P] += pLb;
    Lb[S_T] += TLb;
    Lb[S_U] += uLb;
    Lb[S_V] += vLb;
    Lb[S_W] += wLb;
    Rb[S_RHO] += rRb;
    Rb[S_P] += pRb;
    Rb[S_T] += TRb;
    Rb[S_U] += uRb;
    Rb[S_V] += vRb;
    Rb[S_W] += wRb;
}

// thermo_update reverse: bars of (rho, p, T, e) -> bars of the two
// reconstructed variables (a, b) of the thermo pair (gas_model.h:91-116)
__host__ __device__ inline void thermo_vjp(int mode, double R, double Cv, double rho, double p,
                                           double T, double e, const double* sb, double* ab,
                                           double* bb) {
    (void)e;
    double rb = sb[S_RHO], pb = sb[S_P], Tb = sb[S_T], eb = sb[S_E];
    switch (mode) {
        case 0:  // rho_p: T = p / (rho R), e = Cv T; (a, b) = (p, rho)
            Tb += Cv * eb;
            pb += Tb / (rho * R);
            rb -= Tb * T / rho;
            *ab = pb;
            *bb = rb;
            break;
        case 1:  // rho_T: p = rho R T, e = Cv T; (a, b) = (rho, T)
            Tb += Cv * eb + pb * rho * R;
            rb += pb * R * T;
            *ab = rb;
            *bb = Tb;
            break;
        case 2:  // rho_u: T = e / Cv, p = rho R T; (a, b) = (rho, e)
            Tb += pb * rho * R;
            rb += pb * R * T;
            *ab = rb;
            *bb = eb + Tb / Cv;
            break;
        default:  // p_T: rho = p / (R T), e = Cv T; (a, b) = (p, T)
            Tb += Cv * eb - rb * p / (R * T * T);
            pb += rb / (R * T);
            *ab = pb;
            *bb = Tb;
            break;
    }
}

}  // namespace ibis
#endif
