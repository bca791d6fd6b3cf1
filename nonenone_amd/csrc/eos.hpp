// eos.hpp — two-constant cubic EOS (PR / PR78 / Soreide-Whitson) for one phase.
//
// Follows models/include/natural/thermo/cubic_eos.hpp:
//   mixingParameters   :252-321   (alpha, A_i, B_i, A_ij = sqrt(A_i A_j)(1-k_ij))
//   phaseMixing        :337-363   (S_i, A, B)
//   cubicCoefficients  :366-376, solveCubic :378-427 (trigonometric / Cardano)
//   liquidRoot/vaporRoot :430-474 (smallest / largest admissible Z > B)
//   fugacityAtRoot     :573-605
// and thermo/soreide_whitson.hpp:26-44 (SW water alpha), cubic_eos.hpp:2389-2404
// (SW aqueous H2O-solute k_ij).  Templated on the scalar type so that the same
// code runs in double (flash) and in Dual<N> (Jacobian assembly).
//
// Arithmetic is written in the reference's operation order; the library is
// built with -ffp-contract=off so that +,-,*,/ and sqrt round identically.
#pragma once

#include "common.hpp"
#include "dual.hpp"

namespace nn
{

template <int NC, class S>
struct MixParams
{
    S Aij[NC][NC];
    S Bi[NC];
};

template <int NC, class S>
struct PhaseMix
{
    S Si[NC];
    S A;
    S B;
};

// Error flag carried through one public call (set where the reference throws).
struct Err
{
    int code = kErrNone;
    __host__ __device__ void raise(int c)
    {
        if (code == kErrNone) code = c;
    }
    __host__ __device__ bool bad() const { return code != kErrNone; }
};

__host__ __device__ inline double swAqueousKij(const FluidParams &f, int i, int j)
{
    const int w = f.water;
    if (i != w && j != w)
        return f.kij[i][j];
    if (i == w && j == w)
        return 0.0;
    const int other = (i == w) ? j : i;
    if (f.swAqSet[other])
        return f.swAqKij[other];
    return f.kij[i][j];
}

template <class S>
__host__ __device__ inline S swWaterAlpha(const S &T, double tcw, double salinity)
{
    const S tr = T / tcw;
    const double saltFactor = 1.0 - 0.0103 * ::pow(salinity, 1.1);
    const S sqrtAlpha =
        S(1.0) + 0.4530 * (S(1.0) - tr * saltFactor) + 0.0034 * (pow(tr, -3.0) - S(1.0));
    return sqrtAlpha * sqrtAlpha;
}

template <int NC, class S>
__host__ __device__ void mixingParameters(const FluidParams &f, const S &P, const S &T,
                                          int role, MixParams<NC, S> &out)
{
    S sqrtAi[NC];
    const bool sw = f.model == NN_MODEL_SW;
    for (int i = 0; i < NC; ++i)
    {
        const S pr = P / f.pc[i];
        const S tr = T / f.tc[i];
        S alpha = 1.0;
        if (sw && i == f.water)
        {
            alpha = swWaterAlpha(T, f.tc[i], f.swSalinity);
        }
        else
        {
            const double omega = f.omega[i];
            double m = 0.37464 + 1.54226 * omega - 0.26992 * omega * omega;
            if (f.eosType == 5 && omega > 0.49)
                m = 0.379642 + 1.48503 * omega - 0.164423 * omega * omega +
                    0.016666 * omega * omega * omega;
            const S alphaRoot = S(1.0) + m * (S(1.0) - sqrt(tr));
            alpha = alphaRoot * alphaRoot;
        }
        const S reducedA = f.omegaA * alpha * pr / (tr * tr);
        sqrtAi[i] = sqrt(reducedA);
        out.Bi[i] = f.omegaB * pr / tr;
    }
    const bool aqueous = sw && role == kWater;
    for (int i = 0; i < NC; ++i)
    {
        for (int j = i; j < NC; ++j)
        {
            const double kij = aqueous ? swAqueousKij(f, i, j) : f.kij[i][j];
            const S aij = sqrtAi[i] * sqrtAi[j] * (1.0 - kij);
            out.Aij[i][j] = aij;
            out.Aij[j][i] = aij;
        }
    }
}

template <int NC, class S>
__host__ __device__ void phaseMixing(const S *x, const MixParams<NC, S> &mp, PhaseMix<NC, S> &out)
{
    out.A = 0.0;
    out.B = 0.0;
    for (int i = 0; i < NC; ++i)
    {
        S si = 0.0;
        for (int j = 0; j < NC; ++j)
            si += x[j] * mp.Aij[i][j];
        out.Si[i] = si;
        out.B += x[i] * mp.Bi[i];
    }
    for (int i = 0; i < NC; ++i)
        out.A += x[i] * out.Si[i];
}

// Roots of Z^3 + e2 Z^2 + e1 Z + e0 = 0; non-positive / missing roots are NaN.
template <class S>
__host__ __device__ void solveCubic(const FluidParams &f, const S &A, const S &B, S roots[3])
{
    const double m1 = f.m1, m2 = f.m2;
    const S e0 = -(A * B + m1 * m2 * B * B * (B + 1.0));
    const S e1 = A - (m1 + m2 - m1 * m2) * B * B - (m1 + m2) * B;
    const S e2 = (m1 + m2 - 1.0) * B - 1.0;
    const S &a = e2;
    const S &b = e1;
    const S &c = e0;
    const S Q = (a * a - 3.0 * b) / 9.0;
    const S R = (2.0 * a * a * a - 9.0 * a * b + 27.0 * c) / 54.0;
    const S disc = R * R - Q * Q * Q;
    const double nan = __builtin_nan("");
    roots[0] = S(nan);
    roots[1] = S(nan);
    roots[2] = S(nan);
    if (sval(disc) < 0.0)
    {
        S arg = R / sqrt(Q * Q * Q);
        if (sval(arg) > 1.0)
            arg = S(1.0);
        else if (sval(arg) < -1.0)
            arg = S(-1.0);
        const S theta = acos(arg);
        roots[0] = -2.0 * sqrt(Q) * cos(theta / 3.0) - a / 3.0;
        roots[1] = -2.0 * sqrt(Q) * cos((theta + 2.0 * kPi) / 3.0) - a / 3.0;
        roots[2] = -2.0 * sqrt(Q) * cos((theta - 2.0 * kPi) / 3.0) - a / 3.0;
    }
    else
    {
        S Sv;
        if (sval(R) > 0.0)
            Sv = -pow(abs(R) + sqrt(disc), 1.0 / 3.0);
        else
            Sv = pow(abs(R) + sqrt(disc), 1.0 / 3.0);
        S Tv = 0.0;
        if (::fabs(sval(Sv)) > 0.0)
            Tv = Q / Sv;
        roots[0] = Sv + Tv - a / 3.0;
    }
    for (int k = 0; k < 3; ++k)
        if (finite(sval(roots[k])) && sval(roots[k]) <= 0.0)
            roots[k] = S(nan);
}

template <class S>
__host__ __device__ S selectRoot(const FluidParams &f, const S &A, const S &B, bool liquid, Err &err)
{
    S roots[3];
    solveCubic(f, A, B, roots);
    bool found = false;
    S best = S(__builtin_nan(""));
    for (int k = 0; k < 3; ++k)
    {
        const double z = sval(roots[k]);
        if (!finite(z) || !(z > sval(B)))
            continue;
        if (!found || (liquid ? (z < sval(best)) : (z > sval(best))))
        {
            best = roots[k];
            found = true;
        }
    }
    if (!found)
        err.raise(kErrThermo);
    return best;
}

// lnphi_i at root Z (fugacityAtRoot without the exp/fugacity products).
template <int NC, class S>
__host__ __device__ void lnPhiAtRoot(const FluidParams &f, const MixParams<NC, S> &mp,
                                     const PhaseMix<NC, S> &mix, const S &Z, S lnPhi[NC],
                                     Err &err)
{
    if (!(sval(mix.B) > 0.0) || !(sval(mix.A) > 0.0) || !(sval(Z - mix.B) > 0.0))
    {
        err.raise(kErrThermo);
        for (int i = 0; i < NC; ++i) lnPhi[i] = S(0.0);
        return;
    }
    const double m1 = f.m1, m2 = f.m2;
    const S commonA = -log(Z - mix.B);
    const S logRatio = log((Z + m2 * mix.B) / (Z + m1 * mix.B));
    const S commonB = logRatio * mix.A / ((m1 - m2) * mix.B);
    const S commonC = (Z - 1.0) / mix.B;
    for (int i = 0; i < NC; ++i)
        lnPhi[i] = commonA + commonB * (2.0 * mix.Si[i] / mix.A - mp.Bi[i] / mix.B) +
                   mp.Bi[i] * commonC;
}

// Scalar phase result used by the flash: Z and phi_i (= exp(lnphi_i)).
template <int NC>
struct PhaseThermo
{
    double Z;
    double phi[NC];
};

template <int NC>
__host__ __device__ void phaseThermoFromParams(const FluidParams &f, const MixParams<NC, double> &mp,
                                               int role, const double *x, PhaseThermo<NC> &out,
                                               Err &err)
{
    PhaseMix<NC, double> mix;
    phaseMixing<NC, double>(x, mp, mix);
    const double Z = selectRoot(f, mix.A, mix.B, role != kGas, err);
    out.Z = Z;
    if (err.bad())
    {
        for (int i = 0; i < NC; ++i) out.phi[i] = 1.0;
        return;
    }
    double lnPhi[NC];
    lnPhiAtRoot<NC, double>(f, mp, mix, Z, lnPhi, err);
    for (int i = 0; i < NC; ++i)
        out.phi[i] = ::exp(lnPhi[i]);
}

} // namespace nn
