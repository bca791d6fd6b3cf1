// cpa.hpp — cubic-plus-association (CPA) phase evaluation, E6.
//
// Restates models/include/natural/thermo/cubic_eos.hpp (paths relative to the
// reference checkout):
//   temperature part of the mixing state   :1356-1414 (cache) / :1423-1539
//   simplified radial distribution          :1541-1559
//   association strength Delta_ij           :1561-1582
//   site fractions: analytic symmetric      :1606-1623
//                   single-donor Newton     :1624-1730
//                   general dense Newton +
//                   damped fixed point      :1731-2031 (dense solve :1233-1302)
//   pressure at density (SRK / PR term)     :2050-2078
//   density root (bracket + bisection)      :2090-2178
//   density root with two implicit-function
//   corrections (first derivatives)         :2191-2248
//   fugacity at density                     :2250-2307
//   phase result / compressibility          :2309-2370, :864-882
//
// S is double (flash, phase update) or Dual<N> (AD properties).  Operations
// keep the reference's operand types and order: quantities the reference holds
// as AD scalars (temperature, composition, mixing sums) are S; per-temperature
// constants, which carry no derivatives there, are double here — products and
// sums with them round identically, and every division by a temperature
// quantity keeps an S divisor.
#pragma once

#include "common.hpp"
#include "dual.hpp"
#include "eos.hpp"

namespace nn
{

// Temperature-only part of the CPA mixing state (cpaTemperatureCache_).
template <int NC>
struct CpaTemp
{
    double T = -1.0;
    double ai[NC], bi[NC], aij[NC][NC];
    double tf[NC][NC];   // exp(eps_ij / (R T)) - 1, 0 when the pair does not associate
    double vol[NC][NC];  // beta_ij
};

template <int NC, class S>
struct CpaMix
{
    S Si[NC];
    S a, b;
};

template <int NC, class S>
struct CpaAssoc
{
    S donorX[NC], acceptorX[NC];
    S g, gPrimeOverG, bonded;
};

// Which association solver the fluid needs (cubic_eos.hpp:1306-1353):
// >= 0 a single symmetric associating component (analytic), or
// -2 - k for a single donor-bearing component k, or -1 for none, or
// kCpaGeneral when several components carry donor sites.
constexpr int kCpaGeneral = -1000;
constexpr int kCpaGeneralMaxNc = 4;

__host__ __device__ inline int cpaScheme(const FluidParams &f)
{
    int assoc = -1;
    bool symmetric = true;
    for (int i = 0; i < f.nc; ++i)
    {
        const int d = f.cpaDonors[i], a = f.cpaAcceptors[i];
        if (d == 0 && a == 0)
            continue;
        if (assoc >= 0 || d <= 0 || d != a)
            symmetric = false;
        assoc = i;
    }
    if (symmetric && assoc >= 0)
    {
        for (int i = 0; i < f.nc; ++i)
            if (f.cpaDonors[i] > 0)
                return i;
    }
    int donor = -1;
    for (int i = 0; i < f.nc; ++i)
    {
        if (f.cpaDonors[i] <= 0)
            continue;
        if (donor >= 0)
            return kCpaGeneral;
        donor = i;
    }
    return donor >= 0 ? -2 - donor : -1;
}

// :1356-1414
template <int NC>
__host__ __device__ void cpaTemperature(const FluidParams &f, double T, CpaTemp<NC> &c)
{
    c.T = T;
    double sqrtAi[NC];
    for (int i = 0; i < NC; ++i)
    {
        const double tr = T / f.tc[i];
        const double alphaRoot = 1.0 + f.cpaC1[i] * (1.0 - ::sqrt(tr));
        c.ai[i] = f.cpaA0[i] * alphaRoot * alphaRoot;
        c.bi[i] = f.cpaB[i];
        sqrtAi[i] = ::sqrt(c.ai[i]);
    }
    for (int i = 0; i < NC; ++i)
        for (int j = 0; j < NC; ++j)
        {
            c.aij[i][j] = sqrtAi[i] * sqrtAi[j] * (1.0 - f.kij[i][j]);
            c.tf[i][j] = 0.0;
            c.vol[i][j] = 0.0;
            if (f.cpaDonors[i] <= 0 || f.cpaAcceptors[j] <= 0)
                continue;
            double eps = f.cpaCrossEps[i][j];
            double beta = f.cpaCrossBeta[i][j];
            if (!(eps > 0.0) || !(beta > 0.0))
            {
                if (f.cpaEps[i] > 0.0 && f.cpaEps[j] > 0.0 && f.cpaBeta[i] > 0.0 &&
                    f.cpaBeta[j] > 0.0)
                {
                    eps = 0.5 * (f.cpaEps[i] + f.cpaEps[j]);
                    beta = ::sqrt(f.cpaBeta[i] * f.cpaBeta[j]);
                }
                else
                {
                    eps = 0.0;
                    beta = 0.0;
                }
            }
            if (eps > 0.0 && beta > 0.0)
            {
                c.tf[i][j] = ::exp(eps / (kGasConstant * T)) - 1.0;
                c.vol[i][j] = beta;
            }
        }
}

// Composition part of cpaMixing_ (:1520-1530).
template <int NC, class S>
__host__ __device__ void cpaMixing(const CpaTemp<NC> &c, const S *x, CpaMix<NC, S> &m)
{
    m.a = S(0.0);
    m.b = S(0.0);
    for (int i = 0; i < NC; ++i)
    {
        m.Si[i] = S(0.0);
        for (int j = 0; j < NC; ++j) m.Si[i] += x[j] * c.aij[i][j];
        m.a += x[i] * m.Si[i];
        m.b += x[i] * c.bi[i];
    }
}

// :1561-1582: Delta_ij = g (exp(eps/RT) - 1) b_ij beta_ij
template <int NC, class S>
__host__ __device__ S cpaDelta(const CpaTemp<NC> &c, int i, int j, const S &g)
{
    const double tf = c.tf[i][j];
    const double beta = c.vol[i][j];
    if (!(tf > 0.0) || !(beta > 0.0))
        return S(0.0);
    const double bij = 0.5 * (c.bi[i] + c.bi[j]);
    return g * tf * bij * beta;
}

// Gaussian elimination with partial pivoting (by value magnitude) and back
// substitution on the leading n x n block, the site Newton step's dense solve
// (:1233-1302); false on a (near-)singular pivot or a non-finite solution.
template <int M, class S>
__host__ __device__ bool cpaSolveDense(S (&a)[M][M], S (&b)[M], int n, S (&x)[M])
{
    for (int col = 0; col < n; ++col)
    {
        int piv = col;
        double mag = ::fabs(sval(a[col][col]));
        for (int r = col + 1; r < n; ++r)
        {
            const double v = ::fabs(sval(a[r][col]));
            if (v > mag)
            {
                piv = r;
                mag = v;
            }
        }
        if (!finite(mag) || mag <= 1.0e-18)
            return false;
        if (piv != col)
        {
            for (int j = 0; j < M; ++j)
            {
                const S t = a[piv][j];
                a[piv][j] = a[col][j];
                a[col][j] = t;
            }
            const S t = b[piv];
            b[piv] = b[col];
            b[col] = t;
        }
        const S diag = a[col][col];
        for (int r = col + 1; r < n; ++r)
        {
            const S factor = a[r][col] / diag;
            a[r][col] = S(0.0);
            for (int j = col + 1; j < n; ++j) a[r][j] -= factor * a[col][j];
            b[r] -= factor * b[col];
        }
    }
    for (int r = n - 1; r >= 0; --r)
    {
        S v = b[r];
        for (int c = r + 1; c < n; ++c) v -= a[r][c] * x[c];
        const S d = a[r][r];
        const double dv = ::fabs(sval(d));
        if (!finite(dv) || dv <= 1.0e-18)
            return false;
        x[r] = v / d;
        if (!finite(sval(x[r])))
            return false;
    }
    return true;
}

// General multi-donor / multi-acceptor site fractions (:1731-2031): couplings
// D_ij = rho x_j A_j Delta_ij (donor i, acceptor j) and C_ji = rho x_i D_i
// Delta_ij; one Jacobi sweep from X = 1; dense Newton on the mass-action
// residuals X (1 + sum C X') - 1 with a halving line search that keeps
// 0 < X <= 1 and lowers the residual; damped fixed point from X = 1 when
// Newton fails; thermo error when that does not converge either.
// Out of line: only fluids with several donor-bearing components take this
// path, and its dense Newton would otherwise be inlined into every EOS call of
// the flash, phase-update and assembly kernels of all fluids.
template <int NC, class S>
__host__ __device__ __noinline__ void cpaGeneralSites(const FluidParams &f, const CpaTemp<NC> &c,
                                                      const S &rho, const S *x,
                                                      CpaAssoc<NC, S> &st, Err &err)
{
    constexpr int M = 2 * NC;
    int dU[NC], aU[NC];
    int count = 0;
    for (int i = 0; i < NC; ++i) dU[i] = f.cpaDonors[i] > 0 ? count++ : -1;
    for (int i = 0; i < NC; ++i) aU[i] = f.cpaAcceptors[i] > 0 ? count++ : -1;
    S dC[NC][NC], aC[NC][NC];
    for (int i = 0; i < NC; ++i)
        for (int j = 0; j < NC; ++j)
        {
            dC[i][j] = S(0.0);
            aC[i][j] = S(0.0);
        }
    for (int i = 0; i < NC; ++i)
    {
        if (f.cpaDonors[i] <= 0)
            continue;
        for (int j = 0; j < NC; ++j)
        {
            if (f.cpaAcceptors[j] <= 0)
                continue;
            const S delta = cpaDelta<NC, S>(c, i, j, st.g);
            dC[i][j] = rho * x[j] * double(f.cpaAcceptors[j]) * delta;
            aC[j][i] = rho * x[i] * double(f.cpaDonors[i]) * delta;
        }
    }
    for (int i = 0; i < NC; ++i)
    {
        if (f.cpaDonors[i] > 0)
        {
            S sum = S(0.0);
            for (int j = 0; j < NC; ++j) sum += dC[i][j];
            st.donorX[i] = S(1.0) / (S(1.0) + sum);
        }
        if (f.cpaAcceptors[i] > 0)
        {
            S sum = S(0.0);
            for (int j = 0; j < NC; ++j) sum += aC[i][j];
            st.acceptorX[i] = S(1.0) / (S(1.0) + sum);
        }
    }
    auto residualNorm = [&](const S *dX, const S *aX) {
        double norm = 0.0;
        for (int i = 0; i < NC; ++i)
        {
            if (f.cpaDonors[i] > 0)
            {
                S den = S(1.0);
                for (int j = 0; j < NC; ++j) den += dC[i][j] * aX[j];
                norm = maxd(norm, ::fabs(sval(dX[i] * den - 1.0)));
            }
            if (f.cpaAcceptors[i] > 0)
            {
                S den = S(1.0);
                for (int j = 0; j < NC; ++j) den += aC[i][j] * dX[j];
                norm = maxd(norm, ::fabs(sval(aX[i] * den - 1.0)));
            }
        }
        return norm;
    };
    bool converged = false;
    const int maxNewton = f.cpaMaxIter < 64 ? f.cpaMaxIter : 64;
    for (int it = 0; it < maxNewton; ++it)
    {
        S jac[M][M], rhs[M], inc[M];
        for (int r = 0; r < M; ++r)
        {
            rhs[r] = S(0.0);
            inc[r] = S(0.0);
            for (int k = 0; k < M; ++k) jac[r][k] = S(0.0);
        }
        double norm = 0.0;
        for (int i = 0; i < NC; ++i)
        {
            const int row = dU[i];
            if (row < 0)
                continue;
            S den = S(1.0);
            for (int j = 0; j < NC; ++j)
            {
                den += dC[i][j] * st.acceptorX[j];
                if (aU[j] >= 0)
                    jac[row][aU[j]] = st.donorX[i] * dC[i][j];
            }
            const S residual = st.donorX[i] * den - 1.0;
            rhs[row] = -residual;
            jac[row][row] = den;
            norm = maxd(norm, ::fabs(sval(residual)));
        }
        for (int i = 0; i < NC; ++i)
        {
            const int row = aU[i];
            if (row < 0)
                continue;
            S den = S(1.0);
            for (int j = 0; j < NC; ++j)
            {
                den += aC[i][j] * st.donorX[j];
                if (dU[j] >= 0)
                    jac[row][dU[j]] = st.acceptorX[i] * aC[i][j];
            }
            const S residual = st.acceptorX[i] * den - 1.0;
            rhs[row] = -residual;
            jac[row][row] = den;
            norm = maxd(norm, ::fabs(sval(residual)));
        }
        if (norm <= f.cpaTol)
        {
            converged = true;
            break;
        }
        if (!cpaSolveDense<M, S>(jac, rhs, count, inc))
            break;
        bool accepted = false;
        double damping = 1.0;
        for (int ls = 0; ls < 20; ++ls)
        {
            S cd[NC], ca[NC];
            bool physical = true;
            for (int i = 0; i < NC; ++i)
            {
                cd[i] = st.donorX[i];
                ca[i] = st.acceptorX[i];
                if (dU[i] >= 0)
                {
                    cd[i] += damping * inc[dU[i]];
                    const double v = sval(cd[i]);
                    physical = physical && finite(v) && v > 0.0 && v <= 1.0;
                }
                if (aU[i] >= 0)
                {
                    ca[i] += damping * inc[aU[i]];
                    const double v = sval(ca[i]);
                    physical = physical && finite(v) && v > 0.0 && v <= 1.0;
                }
            }
            if (physical)
            {
                const double cn = residualNorm(cd, ca);
                if (finite(cn) && cn < norm)
                {
                    for (int i = 0; i < NC; ++i)
                    {
                        st.donorX[i] = cd[i];
                        st.acceptorX[i] = ca[i];
                    }
                    accepted = true;
                    if (cn <= f.cpaTol)
                        converged = true;
                    break;
                }
            }
            damping *= 0.5;
        }
        if (converged || !accepted)
            break;
    }
    if (converged)
        return;
    for (int i = 0; i < NC; ++i)
    {
        st.donorX[i] = S(1.0);
        st.acceptorX[i] = S(1.0);
    }
    for (int it = 0; it < f.cpaMaxIter; ++it)
    {
        S nd[NC], na[NC];
        double change = 0.0;
        for (int i = 0; i < NC; ++i)
        {
            nd[i] = st.donorX[i];
            na[i] = st.acceptorX[i];
            if (f.cpaDonors[i] > 0)
            {
                S sum = S(0.0);
                for (int j = 0; j < NC; ++j) sum += dC[i][j] * st.acceptorX[j];
                const S target = S(1.0) / (S(1.0) + sum);
                nd[i] = st.donorX[i] + f.cpaDamping * (target - st.donorX[i]);
                change = maxd(change, ::fabs(sval(nd[i] - st.donorX[i])));
            }
            if (f.cpaAcceptors[i] > 0)
            {
                S sum = S(0.0);
                for (int j = 0; j < NC; ++j) sum += aC[i][j] * st.donorX[j];
                const S target = S(1.0) / (S(1.0) + sum);
                na[i] = st.acceptorX[i] + f.cpaDamping * (target - st.acceptorX[i]);
                change = maxd(change, ::fabs(sval(na[i] - st.acceptorX[i])));
            }
        }
        for (int i = 0; i < NC; ++i)
        {
            st.donorX[i] = nd[i];
            st.acceptorX[i] = na[i];
        }
        if (change <= f.cpaTol)
            return;
        if (it + 1 == f.cpaMaxIter)
        {
            err.raise(kErrThermo);  // "CPA association-site iteration did not converge."
            return;
        }
    }
}

// :1584-2048 (analytic symmetric, single-donor and general solvers)
template <int NC, class S>
__host__ __device__ void cpaAssociation(const FluidParams &f, const CpaTemp<NC> &c, const S &rho,
                                        const S *x, const CpaMix<NC, S> &m,
                                        CpaAssoc<NC, S> &st, Err &err)
{
    for (int i = 0; i < NC; ++i)
    {
        st.donorX[i] = S(1.0);
        st.acceptorX[i] = S(1.0);
    }
    const S eta = m.b * rho / 4.0;
    // simplified radial distribution g = 1 / (1 - 1.9 eta) (:1541-1549)
    const S den = S(1.0) - 1.9 * eta;
    st.g = S(1.0) / den;
    st.gPrimeOverG = 1.9 * st.g;

    const int scheme = cpaScheme(f);
    if (scheme >= 0)
    {
        const int k = scheme;
        const int sites = f.cpaDonors[k];
        const S delta = cpaDelta<NC, S>(c, k, k, st.g);
        const S A = rho * x[k] * double(sites) * delta;
        const S X = S(2.0) / (S(1.0) + sqrt(S(1.0) + S(4.0) * A));
        st.donorX[k] = X;
        st.acceptorX[k] = X;
    }
    else if (scheme == kCpaGeneral)
    {
        // Built for up to kCpaGeneralMaxNc components: beyond that the dense
        // AD Jacobian of the site Newton outgrows a lane's stack (contexts
        // reject such fluids).
        if constexpr (NC <= kCpaGeneralMaxNc)
            cpaGeneralSites<NC, S>(f, c, rho, x, st, err);
        else
            err.raise(kErrInvalid);
        if (err.bad())
            return;
    }
    else if (scheme <= -2)
    {
        const int donor = -2 - scheme;
        S aC[NC], bC[NC];
        S aSum = S(0.0);
        for (int j = 0; j < NC; ++j)
        {
            aC[j] = S(0.0);
            bC[j] = S(0.0);
            if (f.cpaAcceptors[j] <= 0)
                continue;
            const S delta = cpaDelta<NC, S>(c, donor, j, st.g);
            aC[j] = rho * x[j] * double(f.cpaAcceptors[j]) * delta;
            bC[j] = rho * x[donor] * double(f.cpaDonors[donor]) * delta;
            aSum += aC[j];
        }
        S donorX = S(1.0) / (S(1.0) + aSum);
        double lower = 0.0, upper = 1.0;
        bool converged = false;
        const int maxIt = f.cpaMaxIter < 64 ? f.cpaMaxIter : 64;
        for (int it = 0; it < maxIt; ++it)
        {
            S sum = S(0.0), slope = S(0.0);
            for (int j = 0; j < NC; ++j)
            {
                if (f.cpaAcceptors[j] <= 0)
                    continue;
                const S d = S(1.0) + bC[j] * donorX;
                sum += aC[j] / d;
                slope -= aC[j] * bC[j] / (d * d);
            }
            const S residual = donorX * (S(1.0) + sum) - 1.0;
            const S derivative = S(1.0) + sum + donorX * slope;
            const double rv = sval(residual);
            const double dv = sval(donorX);
            if (rv > 0.0)
                upper = dv;
            else
                lower = dv;
            if (::fabs(rv) <= f.cpaTol)
            {
                converged = true;
                break;
            }
            const double derv = sval(derivative);
            S candidate = donorX;
            if (finite(derv) && ::fabs(derv) > 1.0e-14)
                candidate = donorX - residual / derivative;
            const double cv = sval(candidate);
            if (!finite(cv) || !(cv > lower && cv < upper))
            {
                const double mid = 0.5 * (lower + upper);
                candidate = donorX + (mid - dv);
            }
            donorX = candidate;
        }
        if (!converged)
        {
            err.raise(kErrThermo);
            return;
        }
        st.donorX[donor] = donorX;
        for (int j = 0; j < NC; ++j)
            if (f.cpaAcceptors[j] > 0)
                st.acceptorX[j] = S(1.0) / (S(1.0) + bC[j] * donorX);
    }
    st.bonded = S(0.0);
    for (int i = 0; i < NC; ++i)
        st.bonded += x[i] * (double(f.cpaDonors[i]) * (S(1.0) - st.donorX[i]) +
                             double(f.cpaAcceptors[i]) * (S(1.0) - st.acceptorX[i]));
}

// :2050-2078
template <int NC, class S>
__host__ __device__ S cpaPressure(const FluidParams &f, const CpaTemp<NC> &c, const S &rho,
                                  const S &T, const S *x, const CpaMix<NC, S> &m, Err &err)
{
    if (!(sval(S(1.0) - m.b * rho) > 0.0))
    {
        err.raise(kErrThermo);
        return S(0.0);
    }
    const S q = m.b * rho;
    S attraction = S(1.0) + q;
    if (f.cpaTerm == 1)
    {
        constexpr double sqrtTwo = 1.4142135623730950488;
        attraction = (S(1.0) + (1.0 + sqrtTwo) * q) * (S(1.0) + (1.0 - sqrtTwo) * q);
    }
    const S cubic = kGasConstant * T * rho / (S(1.0) - q) - m.a * rho * rho / attraction;
    CpaAssoc<NC, S> as;
    cpaAssociation<NC, S>(f, c, rho, x, m, as, err);
    if (err.bad())
        return S(0.0);
    const S factor = S(1.0) + as.gPrimeOverG * m.b * rho / 4.0;
    const S association = -0.5 * kGasConstant * T * rho * factor * as.bonded;
    return cubic + association;
}

// Certified outcome of one bracket or bisection evaluation of the density root
// for the single-donor scheme, without the exact site solve.
//
// The exact evaluation r(rho) = cpaPressure(rho) - P solves the donor site
// balance h(X) = X (1 + sum_j aC_j / (1 + bC_j X)) - 1 = 0 by Newton from
// 1 / (1 + sum aC) until the computed |h| <= cpaTol.  With aC, bC >= 0,
// h'(X) = 1 + sum_j aC_j / (1 + bC_j X)^2 >= 1 for X >= 0, so any X lies within
// |h(X)| of the root: the exact solver's X and a warm-started X_a differ by at
// most cpaTol + |h(X_a)| plus the rounding of both computed residuals.  The
// pressure depends on X only through the bonded-site sum
//   x_d n_d (1 - X) + sum_j x_j n_j (1 - 1 / (1 + bC_j X)),
// whose X-derivative is bounded by x_d n_d + sum_j x_j n_j bC_j; that, times
// R T rho (1 + rho b g'/(4 g)) / 2, plus a rounding allowance of 64 ulps on every
// term, bounds |r_exact - r_a| by B/2.  If |r_a| > B + stop, the exact
// evaluation has r_a's sign and does not meet the bisection's |r| <= stop test:
// the search takes the same branch without it.
//
// Returns +1 or -1 (certified sign) or 0 (evaluate exactly).  X carries the
// warm start from one call to the next (outside (0, 1]: none yet).  The
// constants of one root search are taken once (CpaSkip); products regrouped
// against the exact evaluation differ from it by a few ulps, inside the bound's
// 64-ulp allowances.
template <int NC>
struct CpaSkip
{
    int donor = -1;   // the donor-bearing component, -1: no certified skips
    double xd, nd;    // its mole fraction and donor-site count
    double RT;
    double na[NC];    // acceptor-site counts (0: not an acceptor of the donor)
    double kDelta[NC];// Delta_dj / g
    double slopeBound;// without the bC_j terms (x_d n_d)
    double sites;     // sum_j x_j (donor + acceptor sites)
};

template <int NC>
__host__ __device__ void cpaSkipSetup(const FluidParams &f, const CpaTemp<NC> &c, double T,
                                      const double *x, int donor, CpaSkip<NC> &k)
{
    k.donor = donor;
    k.xd = x[donor];
    k.nd = double(f.cpaDonors[donor]);
    k.RT = kGasConstant * T;
    k.slopeBound = k.xd * k.nd;
    k.sites = 0.0;
    for (int j = 0; j < NC; ++j)
    {
        k.sites += x[j] * double(f.cpaDonors[j] + f.cpaAcceptors[j]);
        k.na[j] = f.cpaAcceptors[j] > 0 ? double(f.cpaAcceptors[j]) : 0.0;
        k.kDelta[j] = cpaDelta<NC, double>(c, donor, j, 1.0);
        if (!(k.kDelta[j] > 0.0))
            k.na[j] = 0.0;
    }
}

template <int NC>
__host__ __device__ int cpaCertifiedSign(const FluidParams &f, double rho, const double *x,
                                         const CpaMix<NC, double> &m, double P, double stop,
                                         const CpaSkip<NC> &k, double &X)
{
    constexpr double ulp64 = 64.0 * 1.1102230246251565e-16;
    const double q = m.b * rho;
    double attraction = 1.0 + q;
    if (f.cpaTerm == 1)
    {
        constexpr double sqrtTwo = 1.4142135623730950488;
        attraction = (1.0 + (1.0 + sqrtTwo) * q) * (1.0 + (1.0 - sqrtTwo) * q);
    }
    const double repulsive = k.RT * rho / (1.0 - q);
    const double attractive = m.a * rho * rho / attraction;
    const double cubic = repulsive - attractive;
    const double g = 1.0 / (1.0 - 1.9 * (m.b * rho / 4.0));
    double aC[NC], bC[NC];
    double aSum = 0.0, slopeBound = k.slopeBound;
    for (int j = 0; j < NC; ++j)
    {
        const double delta = g * k.kDelta[j];
        aC[j] = rho * x[j] * k.na[j] * delta;
        bC[j] = rho * k.xd * k.nd * delta;
        aSum += aC[j];
        slopeBound += x[j] * k.na[j] * bC[j];
    }
    if (!(X > 0.0 && X <= 1.0))
        X = 1.0 / (1.0 + aSum);
    const double scale = 0.5 * k.RT * rho * (1.0 + 1.9 * g * m.b * rho / 4.0);
    const double roundCubic = ulp64 * (::fabs(repulsive) + ::fabs(attractive) + ::fabs(P));
    for (int step = 0; step < 3; ++step)
    {
        double sum = 0.0, slope = 0.0, bonded = k.xd * k.nd * (1.0 - X);
        for (int j = 0; j < NC; ++j)
        {
            if (k.na[j] == 0.0)
                continue;
            const double inv = 1.0 / (1.0 + bC[j] * X);
            const double t = aC[j] * inv;
            sum += t;
            slope -= t * bC[j] * inv;
            bonded += x[j] * k.na[j] * (1.0 - inv);
        }
        const double h = X * (1.0 + sum) - 1.0;
        const double association = -scale * bonded;
        const double r = cubic + association - P;
        const double hRound = ulp64 * (1.0 + X * (1.0 + sum));
        const double dX = f.cpaTol + ::fabs(h) + 2.0 * hRound;
        const double bound = 2.0 * (scale * (slopeBound * dX + ulp64 * k.sites) + roundCubic +
                                    ulp64 * ::fabs(association));
        if (!finite(r) || !finite(bound))
            return 0;
        if (::fabs(r) > bound + stop)
            return r > 0.0 ? 1 : -1;
        const double next = X - h / (1.0 + sum + X * slope);
        if (!(next > 0.0 && next <= 1.0))
            return 0;
        X = next;
    }
    return 0;
}

#ifndef NN_CPA_CERTIFIED_SKIP
#define NN_CPA_CERTIFIED_SKIP 1
#endif

// :2090-2178 (scalar density root: bracket, then up to 100 bisections).  With
// Skip, evaluations whose branch cpaCertifiedSign settles are not made; the
// bracket, the midpoints and the returned root are those of the plain search.
template <int NC, bool Skip = (NN_CPA_CERTIFIED_SKIP != 0)>
__host__ __device__ double cpaDensityRootValue(const FluidParams &f, const CpaTemp<NC> &c,
                                               double P, double T, const double *x, bool liquid,
                                               const CpaMix<NC, double> &m, Err &err)
{
    const double b = m.b;
    if (!(b > 0.0))
    {
        err.raise(kErrThermo);
        return 0.0;
    }
    const double rhoMax = 0.999999 / b;
    const double stop = 1.0e-10 * maxd(P, 1.0);
    CpaSkip<NC> skip;
    if constexpr (Skip)
    {
        // single donor-bearing component, a site solver that cannot run out of
        // iterations, non-negative composition (the bound needs aC, bC >= 0)
        const int scheme = cpaScheme(f);
        bool ok = scheme <= -2 && scheme != kCpaGeneral && f.cpaMaxIter >= 32 && f.cpaTol > 0.0;
        for (int i = 0; i < NC; ++i)
            ok = ok && x[i] >= 0.0;
        if (ok)
            cpaSkipSetup<NC>(f, c, T, x, -2 - scheme, skip);
    }
    double warmX = -1.0;
    // The residual, or +-inf where its sign is certified and |residual| > stop
    // (every use below reads only its sign and that test).
    auto residual = [&](double rho) {
        if (skip.donor >= 0)
        {
            const int s = cpaCertifiedSign<NC>(f, rho, x, m, P, stop, skip, warmX);
            if (s != 0)
                return s > 0 ? __builtin_inf() : -__builtin_inf();
        }
        return cpaPressure<NC, double>(f, c, rho, T, x, m, err) - P;
    };
    double lo = 0.0, hi = 0.0, flo = -P, fhi = 0.0;
    if (!liquid)
    {
        hi = mind(rhoMax, maxd(1.0, 0.25 * P / (kGasConstant * T)));
        fhi = residual(hi);
        if (err.bad())
            return 0.0;
        while (fhi < 0.0 && hi < 0.999 * rhoMax)
        {
            lo = hi;
            flo = fhi;
            hi = mind(0.999999 * rhoMax, hi * 1.6 + 1.0);
            fhi = residual(hi);
            if (err.bad())
                return 0.0;
        }
    }
    else
    {
        hi = rhoMax;
        fhi = residual(hi);
        if (err.bad())
            return 0.0;
        lo = 0.82 * hi;
        flo = residual(lo);
        if (err.bad())
            return 0.0;
        for (int i = 0; i < 160 && flo > 0.0; ++i)
        {
            hi = lo;
            fhi = flo;
            lo *= 0.82;
            flo = residual(lo);
            if (err.bad())
                return 0.0;
        }
    }
    if (!(flo <= 0.0 && fhi >= 0.0))
    {
        err.raise(kErrThermo);
        return 0.0;
    }
    for (int it = 0; it < 100; ++it)
    {
        const double mid = 0.5 * (lo + hi);
        const double fm = residual(mid);
        if (err.bad())
            return 0.0;
        if (fm > 0.0)
        {
            hi = mid;
            fhi = fm;
        }
        else
        {
            lo = mid;
            flo = fm;
        }
        if (::fabs(fm) <= 1.0e-10 * maxd(P, 1.0) || (hi - lo) <= 1.0e-12 * maxd(mid, 1.0))
            break;
    }
    return 0.5 * (lo + hi);
}

// cpaDensityRootValue for one state held by a whole wavefront (every lane has
// the same arguments): the same bracket points and bisection midpoints, with
// the pressure evaluations of up to 64 bracket points, or of a 6-level
// bisection tree (63 midpoints), done at once on separate lanes and then
// consumed in the serial order with the serial stopping rules.  Points the
// serial search would not have reached are evaluated but never used (their
// errors included), so root and errors are those of the serial search.
template <int NC>
__host__ __device__ double cpaDensityRootValueCoop(const FluidParams &f, const CpaTemp<NC> &c,
                                                   double P, double T, const double *x,
                                                   bool liquid, const CpaMix<NC, double> &m,
                                                   Err &err)
{
    const double b = m.b;
    if (!(b > 0.0))
    {
        err.raise(kErrThermo);
        return 0.0;
    }
    const double rhoMax = 0.999999 / b;
    const int lane = waveLane();
    auto evalAt = [&](double rho, double &fv, int &code) {
        Err e;
        fv = cpaPressure<NC, double>(f, c, rho, T, x, m, e) - P;
        code = e.code;
    };
    double lo = 0.0, hi = 0.0, flo = -P, fhi = 0.0;
    if (!liquid)
    {
        // points h_0 = min(rhoMax, max(1, P / (4RT))), h_{k+1} = min(0.999999 rhoMax, 1.6 h_k + 1)
        double base = mind(rhoMax, maxd(1.0, 0.25 * P / (kGasConstant * T)));
        bool first = true;
        for (;;)
        {
            double h = base;
            for (int j = 0; j < lane; ++j) h = mind(0.999999 * rhoMax, h * 1.6 + 1.0);
            double fv;
            int code;
            evalAt(h, fv, code);
            bool done = false;
            double last = h;
            for (int j = 0; j < 64 && !done; ++j)
            {
                const int e = waveShfl(code, j);
                if (e != kErrNone)
                {
                    err.raise(e);
                    return 0.0;
                }
                const double hj = waveShfl(h, j), fj = waveShfl(fv, j);
                if (!first)
                {
                    lo = hi;
                    flo = fhi;
                }
                first = false;
                hi = hj;
                fhi = fj;
                last = hj;
                done = !(fhi < 0.0 && hi < 0.999 * rhoMax);
            }
            if (done)
                break;
            base = mind(0.999999 * rhoMax, last * 1.6 + 1.0);
        }
    }
    else
    {
        // points rhoMax, 0.82 rhoMax, then repeated * 0.82; at most 160 shrink steps
        double start = rhoMax;
        int consumed = 0;  // points used so far (0 = none)
        int shrinks = 0;
        bool done = false;
        while (!done)
        {
            double pnt = start;
            for (int j = 0; j < lane; ++j) pnt = (consumed + j == 0) ? 0.82 * pnt : pnt * 0.82;
            double fv;
            int code;
            evalAt(pnt, fv, code);
            double last = pnt;
            for (int j = 0; j < 64 && !done; ++j)
            {
                const int e = waveShfl(code, j);
                if (e != kErrNone)
                {
                    err.raise(e);
                    return 0.0;
                }
                const double pj = waveShfl(pnt, j), fj = waveShfl(fv, j);
                last = pj;
                const int index = consumed + j;
                if (index == 0)
                {
                    hi = pj;
                    fhi = fj;
                }
                else if (index == 1)
                {
                    lo = pj;
                    flo = fj;
                    done = !(shrinks < 160 && flo > 0.0);
                }
                else
                {
                    hi = lo;
                    fhi = flo;
                    lo = pj;
                    flo = fj;
                    ++shrinks;
                    done = !(shrinks < 160 && flo > 0.0);
                }
            }
            if (done)
                break;
            consumed += 64;
            start = (consumed == 0) ? 0.82 * last : last * 0.82;
        }
    }
    if (!(flo <= 0.0 && fhi >= 0.0))
    {
        err.raise(kErrThermo);
        return 0.0;
    }
    // bisection, six levels per round: lane j < 63 evaluates heap node j
    int it = 0;
    bool stop = false;
    while (it < 100 && !stop)
    {
        double a = lo, bb = hi;
        double fm = 0.0;
        int code = kErrNone;
        if (lane < 63)
        {
            const unsigned node = unsigned(lane) + 1u;  // 1-based heap index
            int depth = 0;
            while ((node >> (depth + 1)) != 0u) ++depth;
            for (int d = depth - 1; d >= 0; --d)
            {
                const double mid = 0.5 * (a + bb);
                if ((node >> d) & 1u)
                    a = mid;  // right child: f(mid) <= 0 moved lo up
                else
                    bb = mid;  // left child: f(mid) > 0 moved hi down
            }
            evalAt(0.5 * (a + bb), fm, code);
        }
        int node = 0;
        for (int level = 0; level < 6 && it < 100; ++level)
        {
            const double mid = 0.5 * (lo + hi);
            const int e = waveShfl(code, node);
            if (e != kErrNone)
            {
                err.raise(e);
                return 0.0;
            }
            const double fmid = waveShfl(fm, node);
            if (fmid > 0.0)
            {
                hi = mid;
                fhi = fmid;
            }
            else
            {
                lo = mid;
                flo = fmid;
            }
            ++it;
            if (::fabs(fmid) <= 1.0e-10 * maxd(P, 1.0) || (hi - lo) <= 1.0e-12 * maxd(mid, 1.0))
            {
                stop = true;
                break;
            }
            node = fmid > 0.0 ? 2 * node + 1 : 2 * node + 2;
        }
    }
    return 0.5 * (lo + hi);
}

// :2191-2248: scalar root, then two implicit-function corrections
// rho -= (p(rho) - P) / (dp/drho)_FD, which also carry first derivatives.
// rootSlot (may be null): {root, FD slope of correction 1, of correction 2}.
// rootSlot[0] > 0 holds those of an earlier call with the same values of P, T
// and x (reused: a later derivative chunk of the same cell); otherwise this
// call stores its own there.
template <int NC, class S, bool Coop = false>
__host__ __device__ S cpaDensityRoot(const FluidParams &f, const CpaTemp<NC> &c, const S &P,
                                     const S &T, const S *x, bool liquid, Err &err,
                                     double *rootSlot = nullptr)
{
    double xs[NC];
    for (int i = 0; i < NC; ++i) xs[i] = sval(x[i]);
    const double pV = sval(P), tV = sval(T);
    CpaMix<NC, double> sm;
    cpaMixing<NC, double>(c, xs, sm);
    const bool cached = rootSlot && rootSlot[0] > 0.0;
    double root;
    if (cached)
        root = rootSlot[0];
    else
    {
        root = Coop ? cpaDensityRootValueCoop<NC>(f, c, pV, tV, xs, liquid, sm, err)
                    : cpaDensityRootValue<NC>(f, c, pV, tV, xs, liquid, sm, err);
    }
    S rho = P;
    if (err.bad())
        return rho;
    rho *= 0.0;
    rho += root;
    CpaMix<NC, S> dm;
    cpaMixing<NC, S>(c, x, dm);
    for (int k = 0; k < 2; ++k)
    {
        double derivative;
        if (cached)
            derivative = rootSlot[1 + k];
        else
        {
            const double rv = sval(rho);
            const double h = maxd(1.0e-4, ::fabs(rv) * 2.0e-6);
            const double lower = maxd(1.0e-12, rv - h);
            const double upper = rv + h;
            const double pLower = cpaPressure<NC, double>(f, c, lower, tV, xs, sm, err);
            const double pUpper = cpaPressure<NC, double>(f, c, upper, tV, xs, sm, err);
            if (err.bad())
                return rho;
            derivative = (pUpper - pLower) / (upper - lower);
            if (rootSlot)
                rootSlot[1 + k] = derivative;
        }
        if (!finite(derivative) || ::fabs(derivative) < 1.0e-12)
        {
            err.raise(kErrThermo);
            return rho;
        }
        const S residual = cpaPressure<NC, S>(f, c, rho, T, x, dm, err) - P;
        if (err.bad())
            return rho;
        rho -= residual / derivative;
    }
    if (rootSlot && !cached)
        rootSlot[0] = root;
    return rho;
}

// :2250-2307: Z and ln(phi) at a density.
template <int NC, class S>
__host__ __device__ void cpaAtDensity(const FluidParams &f, const CpaTemp<NC> &c, const S &P,
                                      const S &T, const S *x, const S &rho, S &Z, S *lnPhi,
                                      Err &err)
{
    CpaMix<NC, S> m;
    cpaMixing<NC, S>(c, x, m);
    CpaAssoc<NC, S> as;
    cpaAssociation<NC, S>(f, c, rho, x, m, as, err);
    if (err.bad())
        return;
    const S q = m.b * rho;
    if (!(sval(S(1.0) - q) > 0.0))
    {
        err.raise(kErrThermo);
        return;
    }
    Z = P / (rho * kGasConstant * T);
    S integral = log(S(1.0) + q);
    S integralDerivative = S(1.0) / (S(1.0) + q);
    if (f.cpaTerm == 1)
    {
        constexpr double sqrtTwo = 1.4142135623730950488;
        constexpr double deltaOne = 1.0 + sqrtTwo;
        constexpr double deltaTwo = 1.0 - sqrtTwo;
        constexpr double deltaDifference = 2.0 * sqrtTwo;
        const S first = S(1.0) + deltaOne * q;
        const S second = S(1.0) + deltaTwo * q;
        integral = log(first / second) / deltaDifference;
        integralDerivative = S(1.0) / (first * second);
    }
    const S logZ = log(Z);
    for (int i = 0; i < NC; ++i)
    {
        const double bi = c.bi[i];
        S muCubic = -log(S(1.0) - q) + rho * bi / (S(1.0) - q);
        muCubic -= ((2.0 * m.Si[i] / m.b - m.a * bi / (m.b * m.b)) * integral +
                    (m.a / m.b) * rho * bi * integralDerivative) /
                   (kGasConstant * T);
        S muAssociation = S(0.0);
        if (f.cpaDonors[i] > 0)
            muAssociation += double(f.cpaDonors[i]) * log(as.donorX[i]);
        if (f.cpaAcceptors[i] > 0)
            muAssociation += double(f.cpaAcceptors[i]) * log(as.acceptorX[i]);
        const S dLnGdNi = as.gPrimeOverG * bi * rho / 4.0;
        muAssociation -= 0.5 * dLnGdNi * as.bonded;
        lnPhi[i] = muCubic + muAssociation - logZ;
    }
}

// :2309-2340 without the Gibbs-minimum root option: the selected root's Z and
// ln(phi).  liquid = role is not gas.
template <int NC, class S, bool Coop = false>
__host__ __device__ void cpaPhase(const FluidParams &f, const CpaTemp<NC> &c, const S &P,
                                  const S &T, const S *x, bool liquid, S &Z, S *lnPhi, Err &err,
                                  double *rootSlot = nullptr)
{
    const S rho = cpaDensityRoot<NC, S, Coop>(f, c, P, T, x, liquid, err, rootSlot);
    if (err.bad())
        return;
    cpaAtDensity<NC, S>(f, c, P, T, x, rho, Z, lnPhi, err);
}

// :864-882 compressibility(): Z of the selected root only.
template <int NC, bool Coop = false>
__host__ __device__ double cpaCompressibility(const FluidParams &f, const CpaTemp<NC> &c, double P,
                                              double T, const double *x, bool liquid, Err &err)
{
    const double rho = cpaDensityRoot<NC, double, Coop>(f, c, P, T, x, liquid, err);
    if (err.bad())
        return 1.0;
    return P / (rho * kGasConstant * T);
}

} // namespace nn
