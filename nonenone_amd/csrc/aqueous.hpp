// aqueous.hpp — property closures of the Water role: IAPWS-IF97 water density,
// the Garcia (2001) CO2 apparent molar volume, and the IAPWS-2008 industrial /
// McBride-Wright (2015) H2O-CO2 viscosities.
//
// Restates (paths relative to the reference checkout):
//   IF97 Region 1 / 2 density, Region 3
//   Helmholtz equation and density solve,
//   region selection                  models/include/natural/thermo/aqueous_volume.hpp:19-303
//   Garcia CO2 apparent volume         aqueous_volume.hpp:306-319
//   IAPWS-Garcia aqueous molar volume  aqueous_volume.hpp:327-441
//   Water-role molar density and the
//   aqueous composition domain         thermo/cubic_eos.hpp:440-545
//   IAPWS-2008 industrial viscosity,
//   McBride-Wright H2O-CO2 viscosity   models/include/natural/properties/aqueous_viscosity.hpp:28-176
//   Water-role viscosity selection     fluid_system.hpp:97-133
// The correlation coefficients are those of the published standards (IAPWS
// R7-97 Tables 2, 25 and 30; IAPWS R12-08 Tables 1-2) and papers (Garcia 2001,
// McBride-Wright et al. 2015), which the reference evaluates term by term; the
// same term order is kept so values round alike.
//
// S is double or Dual<N>.  Where the reference throws, Err is raised
// (kErrThermo).
#pragma once

#include "common.hpp"
#include "dual.hpp"
#include "eos.hpp"

namespace nn
{

// Region 1 (T <= 623.15 K): specific Gibbs energy derivative gamma_pi.
template <class S>
__host__ __device__ S if97Region1Density(const S &P, const S &T, Err &err)
{
    if (!(sval(P) > 0.0) || !(sval(T) >= 273.15) || !(sval(T) <= 623.15))
    {
        err.raise(kErrThermo);
        return S(1.0);
    }
    constexpr int I[34] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 2, 2, 2,
                           2, 2, 3, 3, 3, 4, 4, 4, 5, 8, 8, 21, 23, 29, 30, 31, 32};
    constexpr int J[34] = {-2, -1, 0, 1, 2, 3, 4, 5, -9, -7, -1, 0, 1, 3, -3, 0, 1,
                           3, 17, -4, 0, 6, -5, -2, 10, -8, -11, -6, -29, -31, -38,
                           -39, -40, -41};
    constexpr double n[34] = {
        0.14632971213167, -0.84548187169114, -3.7563603672040, 3.3855169168385,
        -0.95791963387872, 0.15772038513228, -0.016616417199501, 0.00081214629983568,
        0.00028319080123804, -0.00060706301565874, -0.018990068218419, -0.032529748770505,
        -0.021841717175414, -0.000052838357969930, -0.00047184321073267,
        -0.00030001780793026, 0.000047661393906987, -0.0000044141845330846,
        -7.2694996297594e-16, -0.000031679644845054, -0.0000028270797985312,
        -8.5205128120103e-10, -0.0000022425281908000, -6.5171222895601e-7,
        -1.4341729937924e-13, -4.0516996860117e-7, -1.2734301741641e-9,
        -1.7424871230634e-10, -6.8762131295531e-19, 1.4478307828521e-20,
        2.6335781662795e-23, -1.1947622640071e-23, 1.8228094581404e-24,
        -9.3537087292458e-26};
    constexpr double pStar = 16.53e6, tStar = 1386.0, gasConstant = 461.526;
    const S pi = P / pStar;
    const S tau = tStar / T;
    S gammaPi = 0.0;
    for (int k = 0; k < 34; ++k)
    {
        if (I[k] == 0)
            continue;
        gammaPi -= n[k] * double(I[k]) * pow(S(7.1) - pi, double(I[k] - 1)) *
                   pow(tau - S(1.222), double(J[k]));
    }
    const S v = pi * gammaPi * gasConstant * T / P;
    if (!(sval(v) > 0.0) || !finite(sval(v)))
    {
        err.raise(kErrThermo);
        return S(1.0);
    }
    return S(1.0) / v;
}

// Region 2 (vapour and low-density supercritical water): ideal + residual gamma_pi.
template <class S>
__host__ __device__ S if97Region2Density(const S &P, const S &T, Err &err)
{
    if (!(sval(P) > 0.0) || !(sval(P) <= 100.0e6) || !(sval(T) >= 273.15) ||
        !(sval(T) <= 1073.15))
    {
        err.raise(kErrThermo);
        return S(1.0);
    }
    constexpr int I[43] = {1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 4, 4, 4,
                           5, 6, 6, 6, 7, 7, 7, 8, 8, 9, 10, 10, 10, 16, 16, 18, 20,
                           20, 20, 21, 22, 23, 24, 24, 24};
    constexpr int J[43] = {0, 1, 2, 3, 6, 1, 2, 4, 7, 36, 0, 1, 3, 6, 35, 1, 2, 3,
                           7, 3, 16, 35, 0, 11, 25, 8, 36, 13, 4, 10, 14, 29, 50, 57,
                           20, 35, 48, 21, 53, 39, 26, 40, 58};
    constexpr double n[43] = {
        -0.17731742473213e-2, -0.17834862292358e-1, -0.45996013696365e-1,
        -0.57581259083432e-1, -0.50325278727930e-1, -0.33032641670203e-4,
        -0.18948987516315e-3, -0.39392777243355e-2, -0.43797295650573e-1,
        -0.26674547914087e-4, 0.20481737692309e-7, 0.43870667284435e-6,
        -0.32277677238570e-4, -0.15033924542148e-2, -0.40668253562649e-1,
        -0.78847309559367e-9, 0.12790717852285e-7, 0.48225372718507e-6,
        0.22922076337661e-5, -0.16714766451061e-10, -0.21171472321355e-2,
        -0.23895741934104e2, -0.59059564324270e-17, -0.12621808899101e-5,
        -0.38946842435739e-1, 0.11256211360459e-10, -0.82311340897998e1,
        0.19809712802088e-7, 0.10406965210174e-18, -0.10234747095929e-12,
        -0.10018179379511e-8, -0.80882908646985e-10, 0.10693031879409,
        -0.33662250574171, 0.89185845355421e-24, 0.30629316876232e-12,
        -0.42002467698208e-5, -0.59056029685639e-25, 0.37826947613457e-5,
        -0.12768608934681e-14, 0.73087610595061e-28, 0.55414715350778e-16,
        -0.94369707241210e-6};
    constexpr double pStar = 1.0e6, tStar = 540.0, gasConstant = 461.526;
    const S pi = P / pStar;
    const S tau = tStar / T;
    S gammaPi = S(1.0) / pi;
    for (int k = 0; k < 43; ++k)
        gammaPi += n[k] * double(I[k]) * pow(pi, double(I[k] - 1)) *
                   pow(tau - S(0.5), double(J[k]));
    const S v = pi * gammaPi * gasConstant * T / P;
    if (!(sval(v) > 0.0) || !finite(sval(v)))
    {
        err.raise(kErrThermo);
        return S(1.0);
    }
    return S(1.0) / v;
}

// Region 3: delta-derivatives of the reduced Helmholtz energy.
template <class S>
struct If97DeltaDerivatives
{
    S delta, phiDelta, phiDeltaDelta;
};

template <class S>
__host__ __device__ If97DeltaDerivatives<S> if97Region3Derivatives(const S &rho, const S &T)
{
    constexpr int I[39] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 3, 3,
                           3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 6, 6, 6, 7, 8, 9, 9, 10, 10, 11};
    constexpr int J[39] = {0, 1, 2, 7, 10, 12, 23, 2, 6, 15, 17, 0, 2, 6, 7, 22, 26, 0,
                           2, 4, 16, 26, 0, 2, 4, 26, 1, 3, 26, 0, 2, 26, 2, 26, 2, 26,
                           0, 1, 26};
    constexpr double n[39] = {
        -0.15732845290239e2, 0.20944396974307e2, -0.76867707878716e1,
        0.26185947787954e1, -0.28080781148620e1, 0.12053369696517e1,
        -0.84566812812502e-2, -0.12654315477714e1, -0.11524407806681e1,
        0.88521043984318, -0.64207765181607, 0.38493460186671,
        -0.85214708824206, 0.48972281541877e1, -0.30502617256965e1,
        0.39420536879154e-1, 0.12558408424308, -0.27999329698710,
        0.13899799569460e1, -0.20189915023570e1, -0.82147637173963e-2,
        -0.47596035734923, 0.43984074473500e-1, -0.44476435428739,
        0.90572070719733, 0.70522450087967, 0.10770512626332,
        -0.32913623258954, -0.50871062041158, -0.22175400873096e-1,
        0.94260751665092e-1, 0.16436278447961, -0.13503372241348e-1,
        -0.14834345352472e-1, 0.57922953628084e-3, 0.32308904703711e-2,
        0.80964802996215e-4, -0.16557679795037e-3, -0.44923899061815e-4};
    constexpr double rhoCrit = 322.0, tCrit = 647.096, logCoefficient = 1.0658070028513;
    If97DeltaDerivatives<S> r;
    r.delta = rho / rhoCrit;
    const S tau = tCrit / T;
    r.phiDelta = logCoefficient / r.delta;
    r.phiDeltaDelta = -logCoefficient / (r.delta * r.delta);
    for (int k = 0; k < 39; ++k)
    {
        if (I[k] == 0)
            continue;
        r.phiDelta += n[k] * double(I[k]) * pow(r.delta, double(I[k] - 1)) *
                      pow(tau, double(J[k]));
        if (I[k] > 1)
            r.phiDeltaDelta += n[k] * double(I[k]) * double(I[k] - 1) *
                               pow(r.delta, double(I[k] - 2)) * pow(tau, double(J[k]));
    }
    return r;
}

template <class S>
__host__ __device__ S if97Region3Pressure(const S &rho, const S &T)
{
    const If97DeltaDerivatives<S> d = if97Region3Derivatives<S>(rho, T);
    return rho * 461.526 * T * d.delta * d.phiDelta;
}

template <class S>
__host__ __device__ S if97Region3PressureSlope(const S &rho, const S &T)
{
    const If97DeltaDerivatives<S> d = if97Region3Derivatives<S>(rho, T);
    return 461.526 * T * (S(2.0) * d.delta * d.phiDelta + d.delta * d.delta * d.phiDeltaDelta);
}

// Safeguarded scalar Newton for rho(P, T) from 500 kg/m3, then three Newton
// steps in S arithmetic from the converged value (implicit P, T derivatives).
template <class S>
__host__ __device__ S if97Region3Density(const S &P, const S &T, Err &err)
{
    const double p = sval(P), t = sval(T);
    if (!(p > 0.0) || !(p <= 100.0e6) || !(t >= 623.15) || !(t <= 863.15))
    {
        err.raise(kErrThermo);
        return S(1.0);
    }
    double rho = 500.0;
    bool converged = false;
    for (int it = 0; it < 80; ++it)
    {
        const double value = if97Region3Pressure<double>(rho, t) - p;
        const double slope = if97Region3PressureSlope<double>(rho, t);
        if (finite(value) && finite(slope) && slope > 0.0)
            rho = clampd(rho - value / slope, 20.0, 900.0);
        else
            rho = 0.5 * (rho + 500.0);
        if (::fabs(value) <= 1.0e-10 * maxd(p, 1.0))
        {
            converged = true;
            break;
        }
    }
    if (!converged || !(rho > 0.0) || !finite(rho))
    {
        err.raise(kErrThermo);
        return S(1.0);
    }
    S result(rho);
    for (int it = 0; it < 3; ++it)
    {
        const S residual = if97Region3Pressure<S>(result, T) - P;
        result -= residual / if97Region3PressureSlope<S>(result, T);
    }
    if (!(sval(result) > 0.0) || !finite(sval(result)))
        err.raise(kErrThermo);
    return result;
}

// Region selection (aqueous_volume.hpp:270-295).
template <class S>
__host__ __device__ S if97WaterDensity(const S &P, const S &T, Err &err)
{
    const double p = sval(P), t = sval(T);
    if (t <= 623.15)
        return if97Region1Density<S>(P, T, err);
    const double boundary =
        (348.05185628969 - 1.1671859879975 * t + 0.0010192970039326 * t * t) * 1.0e6;
    if (t <= 863.15 && p <= boundary)
        return if97Region2Density<S>(P, T, err);
    if (t <= 863.15 && p <= 100.0e6)
        return if97Region3Density<S>(P, T, err);
    if (t <= 1073.15 && p <= 100.0e6 && t > 863.15)
        return if97Region2Density<S>(P, T, err);
    err.raise(kErrThermo);
    return S(1.0);
}

// Garcia (2001) apparent molar volume of dissolved CO2 [m3/mol].
template <class S>
__host__ __device__ S garciaCo2Volume(const S &T)
{
    const S celsius = T - 273.15;
    const S celsius2 = celsius * celsius;
    const S cm3 = 37.51 - 9.585e-2 * celsius + 8.740e-4 * celsius2 - 5.044e-7 * celsius2 * celsius;
    return cm3 * 1.0e-6;
}

// Sum of the mole fractions outside {water, co2} (co2 < 0: outside {water}).
template <int NC, class S>
__host__ __device__ S aqueousUnsupported(const FluidParams &f, const S *x, int co2)
{
    S sum = 0.0;
    for (int c = 0; c < NC; ++c)
        if (c != f.water && c != co2)
            sum += x[c];
    return sum;
}

// The IAPWS-Garcia closure's composition domain and the IAPWS-2008 closure's
// solute limit (cubic_eos.hpp:449-468); true when no closure is configured.
template <int NC>
__host__ __device__ bool aqueousDomain(const FluidParams &f, const double *x)
{
    if (f.aqViscosity == NN_AQ_VISCOSITY_IAPWS2008)
    {
        const double solute = aqueousUnsupported<NC, double>(f, x, -1);
        if (!finite(solute) || solute > f.aqViscosityMax)
            return false;
    }
    if (!f.aqVolume)
        return true;
    const double unsupported = aqueousUnsupported<NC, double>(f, x, f.co2);
    return finite(unsupported) && unsupported <= f.aqVolumeMax;
}

// IAPWS-Garcia molar volume of the Water role (aqueous_volume.hpp:408-436).
// Out of line, like aqueousViscosity: the phase-update and flash kernels call the
// molar density at many sites, and inlining the IF97 evaluation into each would
// multiply the code of kernels that are already at the register limit; without
// a closure configured only the not-taken call remains.
template <int NC, class S>
__host__ __device__ __noinline__ S aqueousMolarVolume(const FluidParams &f, const S &P, const S &T,
                                         const S *x, Err &err)
{
    if (sval(aqueousUnsupported<NC, S>(f, x, f.co2)) > f.aqVolumeMax)
    {
        err.raise(kErrThermo);
        return S(1.0);
    }
    const S water = if97WaterDensity<S>(P, T, err);
    if (err.bad())
        return S(1.0);
    const S v = x[f.water] * f.waterMolarMass / water + x[f.co2] * garciaCo2Volume<S>(T);
    if (!(sval(v) > 0.0) || !finite(sval(v)))
        err.raise(kErrThermo);
    return v;
}

// Phase molar density (cubic_eos.hpp:473-545): P / (R T Z) without a volume
// model; with the IAPWS-Garcia model every phase takes the translated-volume
// form 1 / (Z R T / P - 0), the Water role the aqueous molar volume.
template <int NC, class S>
__host__ __device__ S phaseMolarDensity(const FluidParams &f, const S &P, const S &T, const S *x,
                                        const S &Z, bool waterRole, Err &err)
{
    if (!f.aqVolume)
        return P / (kGasConstant * T * Z);
    if (waterRole)
        return S(1.0) / aqueousMolarVolume<NC, S>(f, P, T, x, err);
    S shift = 0.0;
    for (int c = 0; c < NC; ++c) shift += x[c] * 0.0;
    const S volume = Z * kGasConstant * T / P - shift;
    if (!(sval(volume) > 0.0) || !finite(sval(volume)))
    {
        err.raise(kErrThermo);
        return S(1.0);
    }
    return S(1.0) / volume;
}

// IAPWS-2008 industrial viscosity of pure water (critical enhancement = 1)
// at the IF97 density (aqueous_viscosity.hpp:44-101), Pa s.
template <class S>
__host__ __device__ S iapws2008Viscosity(const S &density, const S &T, Err &err)
{
    const S reducedT = T / 647.096;
    const S reducedRho = density / 322.0;
    constexpr double h[4] = {1.67752, 2.20462, 0.6366564, -0.241605};
    S denominator = 0.0;
    S power = 1.0;
    for (int i = 0; i < 4; ++i)
    {
        if (i > 0)
            power *= reducedT;
        denominator += h[i] / power;
    }
    const S mu0 = 100.0 * sqrt(reducedT) / denominator;
    constexpr int ci[21] = {0, 1, 2, 3, 0, 1, 2, 3, 5, 0, 1, 2, 3, 4, 0, 1, 0, 3, 4, 3, 5};
    constexpr int cj[21] = {0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 4, 4, 5, 6, 6};
    constexpr double cv[21] = {5.20094e-1, 8.50895e-2, -1.08374, -2.89555e-1, 2.22531e-1,
                               9.99115e-1, 1.88797, 1.26613, 1.20573e-1, -2.81378e-1,
                               -9.06851e-1, -7.72479e-1, -4.89837e-1, -2.57040e-1,
                               1.61913e-1, 2.57399e-1, -3.25372e-2, 6.98452e-2,
                               8.72102e-3, -4.35673e-3, -5.93264e-4};
    const S tVar = 1.0 / reducedT - 1.0;
    const S rhoVar = reducedRho - 1.0;
    S series = 0.0;
    for (int k = 0; k < 21; ++k)
        series += cv[k] * pow(tVar, double(ci[k])) * pow(rhoVar, double(cj[k]));
    const S mu1 = exp(reducedRho * series);
    const S result = mu0 * mu1 * 1.0e-6;
    if (!(sval(result) > 0.0) || !finite(sval(result)))
        err.raise(kErrThermo);
    return result;
}

// Water-role viscosity (fluid_system.hpp:125-131), Pa s.
template <int NC, class S>
__host__ __device__ __noinline__ S aqueousViscosity(const FluidParams &f, const S &P, const S &T, const S *x,
                                       Err &err)
{
    if (f.aqViscosity == NN_AQ_VISCOSITY_IAPWS2008)
    {
        if (sval(aqueousUnsupported<NC, S>(f, x, -1)) > f.aqViscosityMax)
        {
            err.raise(kErrThermo);
            return S(1.0);
        }
        const S density = if97WaterDensity<S>(P, T, err);
        if (err.bad())
            return S(1.0);
        return iapws2008Viscosity<S>(density, T, err);
    }
    // McBride-Wright et al. (2015), H2O-CO2
    if (sval(aqueousUnsupported<NC, S>(f, x, f.co2)) > f.aqViscosityMax)
    {
        err.raise(kErrThermo);
        return S(1.0);
    }
    constexpr double a = -3.705013, b = 0.00289258, c = 3.98950, d = -0.00326;
    constexpr double e1 = 65.55968, e2 = 2.46811, t0 = 141.5, p0 = 1.0e6;
    const S reducedP = P / p0;
    const S shiftedT = T / t0 - 1.0;
    const S logMilli = a + b * reducedP + (c + d * reducedP) / shiftedT +
                       e1 * exp(-e2 * shiftedT) * x[f.co2];
    const S result = exp(logMilli) * 1.0e-3;
    if (!(sval(result) > 0.0) || !finite(sval(result)))
        err.raise(kErrThermo);
    return result;
}

} // namespace nn
