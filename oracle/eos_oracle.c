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
 *   oracle_newton_limit  state/newton_limiter.hpp:18-245, natural_petsc_runtime_state.inc:452-545
 *   oracle_if97_*, oracle_iapws2008_viscosity, oracle_mcbride_wright_viscosity,
 *   oracle_garcia_co2_volume
 *                        thermo/aqueous_volume.hpp:19-319,
 *                        properties/aqueous_viscosity.hpp:28-176 (see the block at
 *                        the end of this file for their pins)
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

/* Planted mistakes for the tests that show the golden checks catch them
 * (test infrastructure; 0 = the restatement as written): 1 drops the
 * b_i (Z - 1) / B term of ln(phi_i), 2 flips the sign of its attraction
 * (logarithmic) term, 3 drops the binary interaction coefficients. */
static int g_plant = 0;
void oracle_plant(int which) { g_plant = which; }

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
            double k = (sw && role == 2) ? sw_kij(f, i, j) : f->kij[i * NN_MAX_NC + j];
            if (g_plant == 3)
                k = 0.0;   /* planted: binary interaction dropped */
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
    const double cB = (g_plant == 2 ? -1.0 : 1.0) *   /* planted: sign of the attraction term */
                      log((Z + f->m2 * B) / (Z + f->m1 * B)) * A / ((f->m1 - f->m2) * B);
    const double cC = (Z - 1.0) / B;
    for (int i = 0; i < f->nc; ++i)
        lnphi[i] = cA + cB * (2.0 * S[i] / A - m.bi[i] / B) + (g_plant == 1 ? 0.0 : m.bi[i] * cC);
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

/* ---------------------------------------------------------------------------
 * Water-role closures (restated from models/include/natural/thermo/
 * aqueous_volume.hpp:19-441 and properties/aqueous_viscosity.hpp:28-176; the
 * coefficients are those of IAPWS R7-97 and R12-08).  Pinned by the IAPWS
 * verification tables: R7-97 Table 5 (Region 1 specific volume), Table 15
 * (Region 2), Table 33 (Region 3 pressure from density); R12-08 Table 4
 * (viscosity without the critical enhancement) — tests/test_oracle.py.
 * ------------------------------------------------------------------------- */

static double if97_region1_volume(double p, double t)
{
    static const int I[34] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 2, 2, 2,
                              2, 2, 3, 3, 3, 4, 4, 4, 5, 8, 8, 21, 23, 29, 30, 31, 32};
    static const int J[34] = {-2, -1, 0, 1, 2, 3, 4, 5, -9, -7, -1, 0, 1, 3, -3, 0, 1,
                              3, 17, -4, 0, 6, -5, -2, 10, -8, -11, -6, -29, -31, -38,
                              -39, -40, -41};
    static const double n[34] = {
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
    const double pi = p / 16.53e6, tau = 1386.0 / t;
    double g = 0.0;
    for (int k = 0; k < 34; ++k)
        if (I[k] != 0)
            g -= n[k] * I[k] * pow(7.1 - pi, I[k] - 1) * pow(tau - 1.222, J[k]);
    return pi * g * 461.526 * t / p;
}

static double if97_region2_volume(double p, double t)
{
    static const int I[43] = {1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 4, 4, 4,
                              5, 6, 6, 6, 7, 7, 7, 8, 8, 9, 10, 10, 10, 16, 16, 18, 20,
                              20, 20, 21, 22, 23, 24, 24, 24};
    static const int J[43] = {0, 1, 2, 3, 6, 1, 2, 4, 7, 36, 0, 1, 3, 6, 35, 1, 2, 3,
                              7, 3, 16, 35, 0, 11, 25, 8, 36, 13, 4, 10, 14, 29, 50, 57,
                              20, 35, 48, 21, 53, 39, 26, 40, 58};
    static const double n[43] = {
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
    const double pi = p / 1.0e6, tau = 540.0 / t;
    double g = 1.0 / pi;
    for (int k = 0; k < 43; ++k)
        g += n[k] * I[k] * pow(pi, I[k] - 1) * pow(tau - 0.5, J[k]);
    return pi * g * 461.526 * t / p;
}

static void if97_region3_phi(double rho, double t, double *delta, double *phid, double *phidd)
{
    static const int I[39] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 3, 3,
                              3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 6, 6, 6, 7, 8, 9, 9, 10, 10, 11};
    static const int J[39] = {0, 1, 2, 7, 10, 12, 23, 2, 6, 15, 17, 0, 2, 6, 7, 22, 26, 0,
                              2, 4, 16, 26, 0, 2, 4, 26, 1, 3, 26, 0, 2, 26, 2, 26, 2, 26,
                              0, 1, 26};
    static const double n[39] = {
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
    const double d = rho / 322.0, tau = 647.096 / t;
    double a = 1.0658070028513 / d, b = -1.0658070028513 / (d * d);
    for (int k = 0; k < 39; ++k)
    {
        if (I[k] == 0)
            continue;
        a += n[k] * I[k] * pow(d, I[k] - 1) * pow(tau, J[k]);
        if (I[k] > 1)
            b += n[k] * I[k] * (I[k] - 1) * pow(d, I[k] - 2) * pow(tau, J[k]);
    }
    *delta = d;
    *phid = a;
    *phidd = b;
}

/* Region 3 pressure [Pa] from density [kg/m3] and temperature [K]. */
double oracle_if97_region3_pressure(double rho, double t)
{
    double d, a, b;
    if97_region3_phi(rho, t, &d, &a, &b);
    return rho * 461.526 * t * d * a;
}

/* IF97 water density [kg/m3] with the closure's region choice and Region-3
 * Newton solve (aqueous_volume.hpp:182-225, 270-295); NaN outside Regions 1-3. */
double oracle_if97_density(double p, double t)
{
    if (t <= 623.15)
        return (p > 0.0 && t >= 273.15) ? 1.0 / if97_region1_volume(p, t) : NAN;
    const double boundary =
        (348.05185628969 - 1.1671859879975 * t + 0.0010192970039326 * t * t) * 1.0e6;
    if ((t <= 863.15 && p <= boundary) || (t > 863.15 && t <= 1073.15 && p <= 100.0e6))
        return (p > 0.0) ? 1.0 / if97_region2_volume(p, t) : NAN;
    if (t > 863.15 || p > 100.0e6 || !(p > 0.0))
        return NAN;
    double rho = 500.0;
    for (int it = 0; it < 80; ++it)
    {
        double d, a, b;
        if97_region3_phi(rho, t, &d, &a, &b);
        const double value = rho * 461.526 * t * d * a - p;
        const double slope = 461.526 * t * (2.0 * d * a + d * d * b);
        if (isfinite(value) && isfinite(slope) && slope > 0.0)
        {
            rho -= value / slope;
            rho = rho < 20.0 ? 20.0 : (rho > 900.0 ? 900.0 : rho);
        }
        else
            rho = 0.5 * (rho + 500.0);
        if (fabs(value) <= 1.0e-10 * (p > 1.0 ? p : 1.0))
            return rho;
    }
    return NAN;
}

/* Specific volume [m3/kg] of Region 1 or 2 (IAPWS R7-97 verification). */
double oracle_if97_volume(int region, double p, double t)
{
    return region == 1 ? if97_region1_volume(p, t) : if97_region2_volume(p, t);
}

/* IAPWS-2008 viscosity [Pa s] at density [kg/m3], critical enhancement 1. */
double oracle_iapws2008_viscosity(double rho, double t)
{
    static const double h[4] = {1.67752, 2.20462, 0.6366564, -0.241605};
    static const int ci[21] = {0, 1, 2, 3, 0, 1, 2, 3, 5, 0, 1, 2, 3, 4, 0, 1, 0, 3, 4, 3, 5};
    static const int cj[21] = {0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 4, 4, 5, 6, 6};
    static const double cv[21] = {5.20094e-1, 8.50895e-2, -1.08374, -2.89555e-1, 2.22531e-1,
                                  9.99115e-1, 1.88797, 1.26613, 1.20573e-1, -2.81378e-1,
                                  -9.06851e-1, -7.72479e-1, -4.89837e-1, -2.57040e-1,
                                  1.61913e-1, 2.57399e-1, -3.25372e-2, 6.98452e-2,
                                  8.72102e-3, -4.35673e-3, -5.93264e-4};
    const double tr = t / 647.096, dr = rho / 322.0;
    double den = 0.0;
    for (int i = 0; i < 4; ++i) den += h[i] / pow(tr, i);
    const double mu0 = 100.0 * sqrt(tr) / den;
    double s = 0.0;
    for (int k = 0; k < 21; ++k) s += cv[k] * pow(1.0 / tr - 1.0, ci[k]) * pow(dr - 1.0, cj[k]);
    return mu0 * exp(dr * s) * 1.0e-6;
}

/* McBride-Wright (2015) H2O-CO2 viscosity [Pa s]. */
double oracle_mcbride_wright_viscosity(double p, double t, double x_co2)
{
    const double pr = p / 1.0e6, ts = t / 141.5 - 1.0;
    return exp(-3.705013 + 0.00289258 * pr + (3.98950 - 0.00326 * pr) / ts +
               65.55968 * exp(-2.46811 * ts) * x_co2) * 1.0e-3;
}

/* Garcia (2001) apparent molar volume of CO2 in water [m3/mol]. */
double oracle_garcia_co2_volume(double t)
{
    const double c = t - 273.15;
    return (37.51 - 9.585e-2 * c + 8.740e-4 * c * c - 5.044e-7 * c * c * c) * 1.0e-6;
}

/* ---------------------------------------------------------------------------
 * Newton-step limiter (restated from models/include/natural/state/
 * newton_limiter.hpp:18-245 and the driver natural_petsc_runtime_state.inc:
 * 452-545) on the fully compositional primary layout of nc components
 * (pressure, three composition blocks of nc-1, three saturations, [well
 * pressure]).  step: PETSc's Y, overwritten with -(limited delta).  var_scale
 * (np values or NULL), max_aq < 0: no aqueous domain limit.  Pinned by closed
 * forms of each limit (tests/test_oracle.py) and the reference build.
 * ------------------------------------------------------------------------- */
static double lim_scale(double magnitude, double maximum)
{
    if (!(magnitude > 0.0) || !isfinite(magnitude))
        return 1.0;
    const double r = maximum / magnitude;
    return r < 1.0 ? r : 1.0;
}

void oracle_newton_limit(int nc, int has_well, int64_t n, const double *primary, double *step,
                         const double *var_scale, double max_aq, int water, int co2)
{
    const int np = 3 * nc + 1 + (has_well ? 1 : 0);
    const int sat = 3 * nc - 2;
    for (int64_t c = 0; c < n; ++c)
    {
        const double *x = primary + c * np;
        double *y = step + c * np;
        double d[3 * NC_MAX + 2];
        for (int q = 0; q < np; ++q)
        {
            d[q] = -y[q];
            if (var_scale)
                d[q] *= var_scale[q];
        }
        double m = fabs(d[sat]);
        for (int p = 1; p < 3; ++p)
            if (m < fabs(d[sat + p]))
                m = fabs(d[sat + p]);
        const double ss = lim_scale(m, 0.1);
        for (int p = 0; p < 3; ++p) d[sat + p] *= ss;
        for (int p = 0; p < 3; ++p)
        {
            const int b = 1 + p * (nc - 1);
            double mx = 0.0, dep = 0.0;
            for (int i = 0; i < nc - 1; ++i)
            {
                if (mx < fabs(d[b + i]))
                    mx = fabs(d[b + i]);
                dep -= d[b + i];
            }
            if (mx < fabs(dep))
                mx = fabs(dep);
            const double s = lim_scale(mx, 0.1);
            for (int i = 0; i < nc - 1; ++i) d[b + i] *= s;
        }
        const double pr = fabs(x[0]);
        if (pr > 1.0e-14)
            d[0] *= lim_scale(fabs(d[0]) / pr, 0.25);
        if (max_aq >= 0.0)
        {
            const int b = 1 + 2 * (nc - 1);
            double u = 0.0, ud = 0.0, dep = 1.0, depd = 0.0;
            for (int i = 0; i < nc - 1; ++i)
            {
                dep -= x[b + i];
                depd -= d[b + i];
                if (i != water && i != co2)
                {
                    u += x[b + i];
                    ud += d[b + i];
                }
            }
            if (nc - 1 != water && nc - 1 != co2)
            {
                u += dep;
                ud += depd;
            }
            if (ud > 0.0)
            {
                const double margin = max_aq * 1.0e-10 > 1.0e-14 ? max_aq * 1.0e-10 : 1.0e-14;
                const double target = max_aq - margin > 0.0 ? max_aq - margin : 0.0;
                double s = (target - u) / ud;
                s = s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s);
                for (int i = 0; i < nc - 1; ++i) d[b + i] *= s;
            }
        }
        for (int q = 0; q < np; ++q) y[q] = -d[q];
    }
}
