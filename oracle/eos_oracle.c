/*
 * ORACLE — test infrastructure only.  A plain-C restatement of the reference's
 * closed-form per-phase thermodynamics; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product library never
 * links it, and nothing in it is a fallback for a missing GPU.
 *
 * Restated functions (paths relative to the reference checkout):
 *   oracle_mixing        models/include/natural/thermo/cubic_eos.hpp:581-650
 *                        (+ thermo/soreide_whitson.hpp:26-44, cubic_eos.hpp:2389-2404)
 *   oracle_phase         cubic_eos.hpp:666-692 (phaseMixing), :695-803 (cubic roots,
 *                        liquid/vapour selection), :902-934 (fugacityAtRoot)
 *   oracle_wilson_k      thermo/detail/three_phase_flash_core.inc:163-176
 *   oracle_rr2           thermo/detail/three_phase_flash_nonlinear.inc:129-187
 *   oracle_lbc_viscosity properties/compositional_properties.hpp:99-152,166-182
 *
 * Pinned by: test/src/unit/natural_core_test.cpp:84-125 golden Z, fugacity,
 * density and LBC viscosity values; the cubic's defining identity
 * Z^3 + e2 Z^2 + e1 Z + e0 = 0 at every returned root; two-component
 * Rachford-Rice closed form; and, where the reference build exists, the
 * reference itself (tests/test_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../include/nonenone_amd.h"

#define NC_MAX NN_MAX_NC

typedef struct
{
    double aij[NC_MAX][NC_MAX];
    double bi[NC_MAX];
} mixing_t;

static double sw_kij(const nn_fluid_desc *f, int i, int j)
{
    const int w = f->water_component;
    if (i != w && j != w)
        return f->kij[i * NN_MAX_NC + j];
    if (i == w && j == w)
        return 0.0;
    const int other = (i == w) ? j : i;
    return f->sw_aqueous_kij_set[other] ? f->sw_aqueous_kij[other] : f->kij[i * NN_MAX_NC + j];
}

/* role: 0 oil, 1 gas, 2 water */
void oracle_mixing(const nn_fluid_desc *f, double P, double T, int role, mixing_t *m)
{
    double sqrtA[NC_MAX];
    const int sw = f->thermo_model == NN_MODEL_SW;
    for (int i = 0; i < f->nc; ++i)
    {
        const double pr = P / f->pc[i];
        const double tr = T / f->tc[i];
        double alpha;
        if (sw && i == f->water_component)
        {
            const double salt = 1.0 - 0.0103 * pow(f->sw_salinity, 1.1);
            const double s = 1.0 + 0.4530 * (1.0 - tr * salt) + 0.0034 * (pow(tr, -3.0) - 1.0);
            alpha = s * s;
        }
        else
        {
            const double w = f->acentric[i];
            double mm = 0.37464 + 1.54226 * w - 0.26992 * w * w;
            if (f->eos_type == 5 && w > 0.49)
                mm = 0.379642 + 1.48503 * w - 0.164423 * w * w + 0.016666 * w * w * w;
            const double r = 1.0 + mm * (1.0 - sqrt(tr));
            alpha = r * r;
        }
        sqrtA[i] = sqrt(f->omega_a * alpha * pr / (tr * tr));
        m->bi[i] = f->omega_b * pr / tr;
    }
    for (int i = 0; i < f->nc; ++i)
        for (int j = i; j < f->nc; ++j)
        {
            const double k = (sw && role == 2) ? sw_kij(f, i, j) : f->kij[i * NN_MAX_NC + j];
            m->aij[i][j] = m->aij[j][i] = sqrtA[i] * sqrtA[j] * (1.0 - k);
        }
}

/* Real positive roots of the PR cubic; returns the count written to z[]. */
static int cubic_roots(double m1, double m2, double A, double B, double z[3])
{
    const double e0 = -(A * B + m1 * m2 * B * B * (B + 1.0));
    const double e1 = A - (m1 + m2 - m1 * m2) * B * B - (m1 + m2) * B;
    const double e2 = (m1 + m2 - 1.0) * B - 1.0;
    const double Q = (e2 * e2 - 3.0 * e1) / 9.0;
    const double R = (2.0 * e2 * e2 * e2 - 9.0 * e2 * e1 + 27.0 * e0) / 54.0;
    const double disc = R * R - Q * Q * Q;
    double r[3];
    int n = 0;
    if (disc < 0.0)
    {
        double arg = R / sqrt(Q * Q * Q);
        if (arg > 1.0) arg = 1.0;
        if (arg < -1.0) arg = -1.0;
        const double th = acos(arg);
        const double pi = 3.141592653589793238462643383279502884;
        r[0] = -2.0 * sqrt(Q) * cos(th / 3.0) - e2 / 3.0;
        r[1] = -2.0 * sqrt(Q) * cos((th + 2.0 * pi) / 3.0) - e2 / 3.0;
        r[2] = -2.0 * sqrt(Q) * cos((th - 2.0 * pi) / 3.0) - e2 / 3.0;
        n = 3;
    }
    else
    {
        double S = pow(fabs(R) + sqrt(disc), 1.0 / 3.0);
        if (R > 0.0) S = -S;
        const double Tq = fabs(S) > 0.0 ? Q / S : 0.0;
        r[0] = S + Tq - e2 / 3.0;
        n = 1;
    }
    int k = 0;
    for (int i = 0; i < n; ++i)
        if (isfinite(r[i]) && r[i] > 0.0)
            z[k++] = r[i];
    return k;
}

/* Phase result: Z and ln(phi_i).  Returns 0 on success, NN_STATUS_THERMO_ERROR
 * when no admissible root exists. */
int oracle_phase(const nn_fluid_desc *f, double P, double T, const double *x, int role,
                 double *Zout, double *lnphi)
{
    mixing_t m;
    oracle_mixing(f, P, T, role, &m);
    double S[NC_MAX], A = 0.0, B = 0.0;
    for (int i = 0; i < f->nc; ++i)
    {
        double s = 0.0;
        for (int j = 0; j < f->nc; ++j) s += x[j] * m.aij[i][j];
        S[i] = s;
        B += x[i] * m.bi[i];
    }
    for (int i = 0; i < f->nc; ++i) A += x[i] * S[i];
    double roots[3];
    const int nr = cubic_roots(f->m1, f->m2, A, B, roots);
    int found = 0;
    double Z = 0.0;
    for (int k = 0; k < nr; ++k)
    {
        if (!(roots[k] > B)) continue;
        if (!found || (role != 1 ? roots[k] < Z : roots[k] > Z)) { Z = roots[k]; found = 1; }
    }
    if (!found || !(A > 0.0) || !(B > 0.0))
        return NN_STATUS_THERMO_ERROR;
    const double cA = -log(Z - B);
    const double cB = log((Z + f->m2 * B) / (Z + f->m1 * B)) * A / ((f->m1 - f->m2) * B);
    const double cC = (Z - 1.0) / B;
    for (int i = 0; i < f->nc; ++i)
        lnphi[i] = cA + cB * (2.0 * S[i] / A - m.bi[i] / B) + m.bi[i] * cC;
    *Zout = Z;
    return 0;
}

void oracle_eos_batch(const nn_fluid_desc *f, int64_t n, const double *P, const double *T,
                      const double *x, const int32_t *role, double *Z, double *lnphi,
                      int32_t *status)
{
    for (int64_t s = 0; s < n; ++s)
        status[s] = oracle_phase(f, P[s], T[s], x + s * f->nc, role[s], Z + s, lnphi + s * f->nc);
}

void oracle_wilson_k(const nn_fluid_desc *f, double P, double T, double *K)
{
    for (int i = 0; i < f->nc; ++i)
    {
        double k = exp(log(f->pc[i] / P) + 5.373 * (1.0 + f->acentric[i]) * (1.0 - f->tc[i] / T));
        K[i] = k < 1e-30 ? 1e-30 : (k > 1e30 ? 1e30 : k);
    }
}

/* Two-phase Rachford-Rice: safeguarded Newton/bisection on beta in [0,1]. */
int oracle_rr2(int nc, const double *z, const double *K, double beta0, double tol, int maxit,
               double *beta_out)
{
    double f0 = 0.0, f1 = 0.0;
    for (int i = 0; i < nc; ++i)
    {
        f0 += z[i] * (K[i] - 1.0);
        f1 += z[i] * (K[i] - 1.0) / K[i];
    }
    if (f0 <= 0.0) { *beta_out = 0.0; return 1; }
    if (f1 >= 0.0) { *beta_out = 1.0; return 1; }
    double lo = 0.0, hi = 1.0, b = beta0 < 0.0 ? 0.0 : (beta0 > 1.0 ? 1.0 : beta0);
    for (int it = 0; it < maxit; ++it)
    {
        double fv = 0.0, df = 0.0;
        for (int i = 0; i < nc; ++i)
        {
            const double a = K[i] - 1.0, D = (1.0 - b) + b * K[i];
            fv += z[i] * a / D;
            df -= z[i] * a * a / (D * D);
        }
        if (fabs(fv) < tol) { *beta_out = b; return 1; }
        if (fv > 0.0) lo = b; else hi = b;
        double c = fabs(df) > 1e-24 ? b - fv / df : b;
        if (!isfinite(c) || c <= lo || c >= hi) c = 0.5 * (lo + hi);
        b = c;
    }
    return 0;
}

/* LBC viscosity of a phase with molar density rho [mol/m3]. */
double oracle_lbc_viscosity(const nn_fluid_desc *f, const double *x, double rho, double T)
{
    const double psi = (0.45359237 * 9.80665) / (0.0254 * 0.0254);
    const double cp = 1.0e-2 * 0.1;
    const double rk = 5.0 / 9.0;
    double num = 0.0, den = 0.0, ppc = 0.0, tpc = 0.0, vpc = 0.0, mw = 0.0;
    for (int i = 0; i < f->nc; ++i)
    {
        const double mws = sqrt(1000.0 * f->molar_mass[i]);
        const double eps = 5.4402 * pow(f->tc[i] / rk, 1.0 / 6.0) /
                           (mws * pow(f->pc[i] / psi, 2.0 / 3.0) * cp);
        const double tr = T / f->tc[i];
        const double mu = tr > 1.5 ? 17.78e-5 * pow(4.58 * tr - 1.67, 0.625) / eps
                                   : 34.0e-5 * pow(tr, 0.94) / eps;
        num += x[i] * mu * mws;
        den += x[i] * mws;
        ppc += f->pc[i] * x[i];
        tpc += f->tc[i] * x[i];
        vpc += f->vc[i] * x[i];
        mw += f->molar_mass[i] * x[i];
    }
    const double epsMix = 5.4402 * pow(tpc / rk, 1.0 / 6.0) /
                          (sqrt(1000.0 * mw) * pow(ppc / psi, 2.0 / 3.0) * cp);
    const double r = vpc * rho;
    const double poly = 0.1023 + 0.023364 * r + 0.058533 * r * r - 0.040758 * r * r * r +
                        0.0093324 * r * r * r * r;
    return num / den + (poly * poly * poly * poly - 1.0e-4) / epsMix;
}
