// Device-side physics, neighbour codes and random numbers of the collision operator (gfx950).
//
// Floating-point contraction is disabled for everything in this header: each expression is
// evaluated in the reference's source order, one rounding per operation, so FP32 distances are
// bit-exact and FP64 pair terms differ from a host evaluation only through library transcendentals.
#ifndef GDX_DEVICE_H
#define GDX_DEVICE_H

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cstdint>

#include "gdx_model.h"

#pragma clang fp contract(off)

// the pair physics also compiles for the host, where a check program compares its hoisted and
// merged forms with the direct per-pair forms bit for bit
#define GDX_FN __host__ __device__ __forceinline__

namespace gdx {

// ---------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). Counter words {draw lo, draw hi, seed hi,
// tag}, key {particle index, seed lo}: the stream of a particle does not depend on storage order.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r)
    {
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        uint32_t lo0 = 0xD2511F53u*c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        uint32_t lo1 = 0xCD9E8D57u*c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// uniform in (0,1) from 52 random bits: (m + 1/2) 2^-52
__device__ __forceinline__ double uniform01(uint32_t pidx, uint64_t seed, uint64_t counter)
{
    uint32_t c[4] = {uint32_t(counter), uint32_t(counter >> 32), uint32_t(seed >> 32), 0x47445843u};
    philox10(c, pidx, uint32_t(seed));
    uint64_t m = (uint64_t(c[1]) << 32) | c[0];
    return (double(m >> 12) + 0.5)*0x1p-52;
}

// draw `counter` of a particle's stream, without advancing the stream
__device__ __forceinline__ double col_uniform_at(uint32_t pidx, uint64_t seed, uint64_t counter)
{
    return fmin(uniform01(pidx, seed, counter), 0x1.fffffffffffffp-1);
}

// _get_col_uniform (inc/swarm/_col_rates.cuh:145-149)
__device__ __forceinline__ double col_uniform(uint32_t pidx, uint64_t seed, uint64_t &counter)
{
    double u = col_uniform_at(pidx, seed, counter);
    counter++;
    return u;
}

// ---------------------------------------------------------------------------------------------
// disk and grain profiles (inc/swarm/param_phys.cuh:16-210, param_grid.cuh:23-27)
// ---------------------------------------------------------------------------------------------
GDX_FN double grain_mass(double size) { return PI*RHO_0*size*size*size / 6.0; }
GDX_FN double omegaK(double R) { return sqrt(G*M_S / R / R / R); }
GDX_FN double hg(double R) { return ASPR_0*pow(R / R_0, 0.5*(IDX_Q + 1.0)); }
GDX_FN double eta(double R, double Z, double h)
{ return -0.5*((IDX_P + 0.5*IDX_Q - 1.5)*h*h + IDX_Q*(1.0 - R / sqrt(R*R + Z*Z))); }
GDX_FN double gas_strat(double R, double Z, double h)
{ return exp((R / sqrt(R*R + Z*Z) - 1.0) / (h*h)); }
GDX_FN double sigma_g(double R) { return SIGMA_0*pow(R / R_0, IDX_P); }
GDX_FN double sound_speed(double R, double h) { return h*omegaK(R)*R; }
GDX_FN double alpha_of(double R, double h)
{
#ifdef CONST_NU
    return NU / (h*h*R*R*omegaK(R));
#else
    (void)R; (void)h;
    return ALPHA;
#endif
}
GDX_FN double cyl_R(double y, double z) { return (N_Z > 1) ? y*sin(z) : y; }
GDX_FN double cyl_Z(double y, double z) { return (N_Z > 1) ? y*cos(z) : 0.0; }

// _get_re_inv_sqrt (inc/swarm/_collision.cuh:136-150)
GDX_FN double re_inv_sqrt(double alpha, double sig)
{
#ifdef CODE_UNIT
    double alpha_0 = alpha_of(R_0, ASPR_0);
    double reynolds = REYNOLDS_0*(alpha / alpha_0)*(sig / SIGMA_0);
#else
    double reynolds = 0.5*alpha*sig*X_SEC / M_MOL;
#endif
    return 1.0 / sqrt(reynolds);
}

// owner environment: every owner-only factor of _get_vrel_pair (_collision.cuh:285-317), in order
GDX_FN gdx_env make_env(double y, double z)
{
    double R = cyl_R(y, z);
    double Z = cyl_Z(y, z);
    double h = hg(R);
    double omega = omegaK(R);
    double strat = gas_strat(R, Z, h);
    double c_s = sound_speed(R, h);
    double alpha = alpha_of(R, h);
    gdx_env e;
    e.Z = Z;
    e.omega = omega;
    e.vn = -eta(R, Z, h)*R*omega;
#ifndef CONST_ST
    e.radial = pow(R / R_0, IDX_P);
#else
    e.radial = 1.0;
#endif
    e.strat = strat;
    e.cs = c_s;
    e.re_inv_sqrt = re_inv_sqrt(alpha, sigma_g(R)*strat);
    e.vg_sq = 1.5*alpha*c_s*c_s;
    return e;
}

// _get_stokes for analytic gas (param_phys.cuh:170-206): divide by the radial and vertical scalings
GDX_FN double stokes_scaled(double size, double radial, double strat)
{
    double st = STOKES_0*(size / S_0);
#ifndef CONST_ST
    st /= radial;
    st /= strat;
#else
    (void)radial; (void)strat;
#endif
    return st;
}
GDX_FN double stokes(const gdx_env &e, double size) { return stokes_scaled(size, e.radial, e.strat); }

// Ormel & Cuzzi (2007) turbulent relative speed, regimes of _get_vrel_t (_collision.cuh:152-273)
GDX_FN double vrel_t(const gdx_env &e, double stokes_i, double stokes_j)
{
    double ris = e.re_inv_sqrt;
    double vg_sq = e.vg_sq;
    double sl, ss;
    if (stokes_i >= stokes_j) { sl = stokes_i; ss = stokes_j; }
    else { sl = stokes_j; ss = stokes_i; }
    double eps = ss / sl;
    const double y_a = 1.6;
    double y_s = 1.6015125;
    y_s += -0.63119577*sl;
    y_s +=  0.32938936*sl*sl;
    y_s += -0.29847604*sl*sl*sl;
    double v;
    if (sl < 0.2*ris)
    {
        v = vg_sq*(sl - ss)*(sl - ss) / ris;
    }
    else if (sl < ris / y_a)
    {
        v = vg_sq*(sl - ss) / (sl + ss);
        v *= (sl / (1.0 + ris / sl) - ss / (1.0 + ris / ss));
    }
    else if (sl < 5.0*ris)
    {
        double coeff = (sl - ss) / (sl + ss);
        coeff *= (sl / (1.0 + y_a) - ss*ss / (ss + y_a*sl));
        coeff += 2.0*(y_a*sl - ris) + sl / (1.0 + y_a);
        coeff -= sl*sl / (sl + ris);
        coeff += ss*ss / (y_a*sl + ss);
        coeff -= ss*ss / (ss + ris);
        v = vg_sq*coeff;
    }
    else if (sl < 0.2)
    {
        v = vg_sq*sl;
        v *= (2.0*y_a - (1.0 + eps) + 2.0 / (1.0 + eps)*(1.0 / (1.0 + y_a) + eps*eps*eps / (y_a + eps)));
    }
    else if (sl < 1.0)
    {
        v = vg_sq*sl;
        v *= (2.0*y_s - (1.0 + eps) + 2.0 / (1.0 + eps)*(1.0 / (1.0 + y_s) + eps*eps*eps / (y_s + eps)));
    }
    else
    {
        v = vg_sq*(1.0 / (1.0 + sl) + 1.0 / (1.0 + ss));
    }
    // a negative square is an invalid state: the NaN propagates to the chain's validity checks
    return v < 0.0 ? __builtin_nan("") : sqrt(v);
}

// vrel_t for the kernels: regimes 3, 4 and 5 share their first three divisions, with operands
// selected per lane (regime 3: (sl - ss)/(sl + ss), sl/(1 + y_a), ss^2/(ss + y_a sl); regimes 4-5:
// eps = ss/sl, 2/(1 + eps), eps^3/(y + eps)), so a wavefront whose lanes straddle sl = 5 Re^-1/2 or
// sl = 0.2 pays two more divisions instead of a second branch. Every value is the same operation on
// the same operands as in vrel_t; ss^2/(y_a sl + ss) equals ss^2/(ss + y_a sl) (addition commutes).
GDX_FN double vrel_t_merged(double ris, double vg_sq, double stokes_i, double stokes_j)
{
    double sl, ss;
    if (stokes_i >= stokes_j) { sl = stokes_i; ss = stokes_j; }
    else { sl = stokes_j; ss = stokes_i; }
    const double y_a = 1.6;
    const bool small = sl < 5.0*ris;
    double v;
    if (small && sl < ris / y_a)
    {
        if (sl < 0.2*ris)
        {
            v = vg_sq*(sl - ss)*(sl - ss) / ris;
        }
        else
        {
            v = vg_sq*(sl - ss) / (sl + ss);
            v *= (sl / (1.0 + ris / sl) - ss / (1.0 + ris / ss));
        }
    }
    else if (small || sl < 1.0)
    {
        const bool r3 = small;
        double y = y_a;
        if (!(sl < 0.2))
        {
            y = 1.6015125;
            y += -0.63119577*sl;
            y +=  0.32938936*sl*sl;
            y += -0.29847604*sl*sl*sl;
        }
        double d1 = (r3 ? sl - ss : ss) / (r3 ? sl + ss : sl);
        double eps = d1;
        double d2 = (r3 ? sl : 2.0) / (r3 ? 1.0 + y_a : 1.0 + eps);
        double d3 = (r3 ? ss*ss : eps*eps*eps) / (r3 ? ss + y_a*sl : y + eps);
        if (r3)
        {
            double coeff = d1;
            coeff *= (d2 - d3);
            coeff += 2.0*(y_a*sl - ris) + d2;
            coeff -= sl*sl / (sl + ris);
            coeff += d3;
            coeff -= ss*ss / (ss + ris);
            v = vg_sq*coeff;
        }
        else
        {
            double inv = 1.0 / (1.0 + y_a);
            if (!(sl < 0.2)) inv = 1.0 / (1.0 + y);
            v = vg_sq*sl;
            v *= (2.0*y - (1.0 + eps) + d2*(inv + d3));
        }
    }
    else
    {
        v = vg_sq*(1.0 / (1.0 + sl) + 1.0 / (1.0 + ss));
    }
    return v < 0.0 ? __builtin_nan("") : sqrt(v);
}

// query-local drift + settling + turbulence (+ Brownian) speed (_collision.cuh:275-325, _col_rates.cuh:127-141)
GDX_FN double vrel_pair(const gdx_env &e, double size_i, double size_j)
{
    double si = stokes(e, size_i);
    double sj = stokes(e, size_j);
    double fi = 1.0 / (1.0 + si*si);
    double fj = 1.0 / (1.0 + sj*sj);
    double dvr = 2.0*e.vn*(si*fi - sj*fj);
    double dvphi = e.vn*(fi - fj);
    double dvz = e.Z*e.omega*(fmin(si, 0.5) - fmin(sj, 0.5));
    double vt = vrel_t(e, si, sj);
    double vb = 0.0;
#ifndef CODE_UNIT
    double mi = grain_mass(size_i);
    double mj = grain_mass(size_j);
    vb = fmin(sqrt(8.0*e.cs*e.cs*M_MOL*(mi + mj) / (PI*mi*mj)), e.cs);
#endif
    return sqrt(dvr*dvr + dvphi*dvphi + dvz*dvz + vt*vt + vb*vb);
}

// KNN ball measure (_collision.cuh:66-115)
GDX_FN double ball_measure(double y, double z, double radius)
{
#ifdef GDX_UNIT_MEASURE
    (void)y; (void)z;
    return radius > 0.0 ? 1.0 : 0.0;
#else
    if (radius <= 0.0) return 0.0;
    if (N_X == 1 && N_Z == 1)
    {
        double R_lo = fmax(Y_MIN, y - radius);
        double R_hi = fmin(Y_MAX, y + radius);
        return (R_hi > R_lo) ? PI*(R_hi*R_hi - R_lo*R_lo) : 0.0;
    }
    constexpr int dim = 1 + int(N_X > 1) + int(N_Z > 1);
    double measure = (dim == 1) ? 2.0*radius : (dim == 2) ? PI*radius*radius
        : 4.0*PI*radius*radius*radius / 3.0;
    double distances[4] = {y - Y_MIN, Y_MAX - y, 1.0e+100, 1.0e+100};
    if (N_Z > 1)
    {
        distances[2] = y*(z - Z_MIN);
        distances[3] = y*(Z_MAX - z);
    }
    constexpr int bounds = (N_Z > 1) ? 4 : 2;
#pragma unroll
    for (int b = 0; b < bounds; b++)
    {
        double d = fmax(0.0, distances[b]);
        if (d >= radius) continue;
        if (dim == 1)
        {
            measure *= 0.5*(1.0 + d / radius);
        }
        else if (dim == 2)
        {
            double cap = radius*radius*acos(d / radius) - d*sqrt(radius*radius - d*d);
            measure *= 1.0 - cap / (PI*radius*radius);
        }
        else
        {
            double cap = PI*(radius - d)*(radius - d)*(2.0*radius + d) / 3.0;
            measure *= 1.0 - cap / (4.0*PI*radius*radius*radius / 3.0);
        }
    }
    if (N_X == 1) measure *= 2.0*PI*cyl_R(y, z);
    return measure;
#endif
}

// ---------------------------------------------------------------------------------------------
// logarithms of the pair outcome moments, after fdlibm's e_log.c / s_log1p.c (Sun Microsystems,
// 1993): 1 + x = 2^k (1 + f) with 1 + f in [sqrt(1/2), sqrt(2)), s = f / (2 + f), and
// log(1 + f) = f - (hfsq - s (hfsq + R(s^2))) with fdlibm's degree-7 minimax polynomial R (Lg1-Lg7).
// Branch-free (every lane runs one instruction stream) and under one ulp; about a third of the
// instructions of a double-double library log1p.
// ---------------------------------------------------------------------------------------------
constexpr double LN2_HI = 6.93147180369123816490e-01;   // 32 significant bits: k LN2_HI is exact
constexpr double LN2_LO = 1.90821492927058770002e-10;
constexpr double LG1 = 6.666666666666735130e-01, LG2 = 3.999999999940941908e-01;
constexpr double LG3 = 2.857142874366239149e-01, LG4 = 2.222219843214978396e-01;
constexpr double LG5 = 1.818357216161805012e-01, LG6 = 1.531383769920937332e-01;
constexpr double LG7 = 1.479819860511658591e-01;
constexpr double SQRT_HALF = 7.07106781186547524401e-01;

// log((1 + f) 2^k) + c for 1 + f in [sqrt(1/2), sqrt(2)) and a small correction c
GDX_FN double log_kernel(double f, int k, double c)
{
    double s = f / (2.0 + f);
    double z = s*s;
    double R = z*fma(z, fma(z, fma(z, fma(z, fma(z, fma(z, LG7, LG6), LG5), LG4), LG3), LG2), LG1);
    double hfsq = 0.5*f*f;
    double dk = static_cast<double>(k);
    return dk*LN2_HI - ((hfsq - (s*(hfsq + R) + (dk*LN2_LO + c))) - f);
}

// 1 + x = 2^k m with m in [sqrt(1/2), sqrt(2)); returns m
GDX_FN double log_reduce(double u, int &k)
{
    double m = frexp(u, &k);
    bool low = m < SQRT_HALF;
    m = low ? 2.0*m : m;
    k = low ? k - 1 : k;
    return m;
}

GDX_FN double ulog1p(double x)
{
    double u = 1.0 + x;
    // u + e = 1 + x exactly (Fast2Sum with the larger operand first)
    double e = (fabs(x) <= 1.0) ? x - (u - 1.0) : 1.0 - (u - x);
    int k;
    double m = log_reduce(u, k);
    // k = 0: f = x exactly; otherwise f = m - 1 (exact) and c ~ e / u, needed to a few bits only
    double r = 2.0 - m;
    r = r*(2.0 - m*r);
    double f = (k == 0) ? x : m - 1.0;
    double c = (k == 0) ? 0.0 : ldexp(e*r, -k);
    double y = log_kernel(f, k, c);
    if (!(x < __builtin_inf())) y = x;                 // NaN, +inf
    if (x == -1.0) y = -__builtin_inf();
    if (x < -1.0) y = __builtin_nan("");
    return y;
}

GDX_FN double ulog(double x)
{
    int k;
    double m = log_reduce(x, k);
    double y = log_kernel(m - 1.0, k, 0.0);
    if (!(x < __builtin_inf())) y = x;                 // NaN, +inf
    if (x == 0.0) y = -__builtin_inf();
    if (x < 0.0) y = __builtin_nan("");
    return y;
}

// ---------------------------------------------------------------------------------------------
// outcomes (inc/swarm/_col_event.cuh:15-109)
// ---------------------------------------------------------------------------------------------
GDX_FN double sticking_packet(double q)
{ return q > 0.0 && q <= 1.0e-06 ? fmax(1.0, floor(1.0e-04 / q)) : 1.0; }
GDX_FN double mass_ratio(double size_i, double size_j)
{ double r = size_j / size_i; return r*r*r; }

GDX_FN double outcome_moments(double si, double sj, bool high_speed,
                                                  double &mean, double &second, double &maximum)
{
    double q = mass_ratio(si, sj);
    double Gp = sticking_packet(q);
    if (!high_speed)
    {
        mean = ulog1p(Gp*q) / 3.0;
        second = mean*mean;
        maximum = mean;
        return 1.0 / Gp;
    }
    if (q <= 0.1)
    {
        double remnant = (1.0 - q) / Gp;
        double debris = q;
        double factor = remnant + debris;
        double jr = -ulog1p(-Gp*q) / 3.0;
        double jd = -ulog(q) / 3.0;
        mean = (remnant*jr + debris*jd) / factor;
        second = (remnant*jr*jr + debris*jd*jd) / factor;
        maximum = fmax(jr, jd);
        return factor;
    }
    double L = 0.5*log(si / INIT_SMIN);
    if (L < 1.0e-03)
    {
        mean = L - L*L / 6.0 + L*L*L*L / 360.0;
        second = L*L*(4.0/3.0 - L / 3.0 + L*L / 90.0 + L*L*L / 180.0);
    }
    else
    {
        double tail = exp(-L) / (-expm1(-L));
        mean = 2.0*(1.0 - L*tail);
        second = 8.0 - (4.0*L*L + 8.0*L)*tail;
    }
    maximum = 2.0*L;
    return 1.0;
}

// outcome category and log mass change of one sampled event (_col_event.cuh:79-109): sticking in four
// mass-ratio bins (0-3), catastrophic fragmentation (4), remnant erosion (5), debris (6)
GDX_FN int outcome_category(double si, double sj, bool high_speed, double u, double size_new, double &log_mass)
{
    double q = mass_ratio(si, sj);
    double Gp = sticking_packet(q);
    if (!high_speed)
    {
        log_mass = log1p(Gp*q);
        return q <= 1.0e-06 ? 0 : q <= 1.0e-04 ? 1 : q <= 1.0e-02 ? 2 : 3;
    }
    if (q <= 0.1)
    {
        double factor = (1.0 - q) / Gp + q;
        if (u < q / factor)
        {
            log_mass = log(q);
            return 6;
        }
        log_mass = log1p(-Gp*q);
        return 5;
    }
    log_mass = 3.0*log(size_new / si);
    return 4;
}

// the new owner size of one event from the pair's mass ratio q = (sj / si)^3
GDX_FN double sample_outcome_q(double si, double sj, double q, bool high_speed, double u)
{
    double Gp = sticking_packet(q);
    if (!high_speed) return cbrt(si*si*si + Gp*sj*sj*sj);
    if (q <= 0.1)
    {
        double factor = (1.0 - q) / Gp + q;
        if (u < q / factor) return sj;
        return si*cbrt(1.0 - Gp*q);
    }
    double lower = sqrt(INIT_SMIN);
    double root = lower + u*(sqrt(si) - lower);
    return fmin(si, fmax(INIT_SMIN, root*root));
}

GDX_FN double sample_outcome(double si, double sj, bool high_speed, double u)
{
    return sample_outcome_q(si, sj, mass_ratio(si, sj), high_speed, u);
}

// RN(a/b) from y = RN(1/b): q = a y, r = a - b q (exact with one FMA), q' = RN(q + r y) is the
// correctly rounded quotient (Markstein's theorem; valid without underflow or overflow). Used where
// the divisor is owner-uniform or a constant, so the reciprocal is paid once.
GDX_FN double div_rn(double a, double b, double y)
{
    double q = a*y;
    double r = fma(-q, b, a);
    return fma(r, y, q);
}

constexpr double INV_S0 = 1.0 / S_0;
constexpr double INV_THREE = 1.0 / 3.0;
constexpr double INV_SMIN = 1.0 / INIT_SMIN;

// owner-only factors of one bath-start or chain evaluation; every value is produced by the same
// operations, in the same order, as the reference's per-pair expressions
struct owner_terms
{
    double size_i, m, inv_m, inv_size;
    double si, fi, ai, mi, vn, zo, ris, vg_sq, radial, inv_radial, strat, inv_strat, cs, hg;
};

// _get_stokes (stokes_scaled) with the owner's scalings: the same three correctly rounded divisions,
// by the held reciprocals
GDX_FN double stokes_partner(const owner_terms &t, double size)
{
    double st = STOKES_0*div_rn(size, S_0, INV_S0);
#ifndef CONST_ST
    st = div_rn(st, t.radial, t.inv_radial);
    st = div_rn(st, t.strat, t.inv_strat);
#endif
    return st;
}

// the owner-size factors, recomputed after each event from the position factors already held
GDX_FN void owner_resize(owner_terms &t, double size_i)
{
    t.size_i = size_i;
    t.inv_size = 1.0 / size_i;
    if constexpr (COAG_KERNEL == 3)
    {
        t.si = stokes_partner(t, size_i);
        t.fi = 1.0 / (1.0 + t.si*t.si);
        t.ai = t.si*t.fi;
        t.mi = fmin(t.si, 0.5);
    }
}

GDX_FN owner_terms make_owner(const gdx_env &e, double size_i, double m, double hg_i)
{
    owner_terms t{};
    t.m = m;
    t.inv_m = 1.0 / m;
    if constexpr (COAG_KERNEL == 3)
    {
        t.vn = e.vn;
        t.zo = e.Z*e.omega;
        t.ris = e.re_inv_sqrt;
        t.vg_sq = e.vg_sq;
        t.radial = e.radial;
        t.inv_radial = 1.0 / e.radial;
        t.strat = e.strat;
        t.inv_strat = 1.0 / e.strat;
        t.cs = e.cs;
        t.hg = hg_i;
    }
    else
    {
        (void)e; (void)hg_i;
    }
    owner_resize(t, size_i);
    return t;
}

// sample_outcome with the mass ratio's division by the owner's size taken through its held reciprocal
// (the correctly rounded quotient, as div_rn gives it): the same size bit for bit
GDX_FN double sample_outcome_owner(const owner_terms &t, double sj, bool high_speed, double u)
{
    double ratio = div_rn(sj, t.size_i, t.inv_size);
    return sample_outcome_q(t.size_i, sj, ratio*ratio*ratio, high_speed, u);
}


GDX_FN double vrel_pair_terms(const owner_terms &t, double size_j)
{
    double sj = stokes_partner(t, size_j);
    double fj = 1.0 / (1.0 + sj*sj);
    double dvr = 2.0*t.vn*(t.ai - sj*fj);
    double dvphi = t.vn*(t.fi - fj);
    double dvz = t.zo*(t.mi - fmin(sj, 0.5));
    double vt = vrel_t_merged(t.ris, t.vg_sq, t.si, sj);
    double vb = 0.0;
#ifndef CODE_UNIT
    double mi = grain_mass(t.size_i);
    double mj = grain_mass(size_j);
    vb = fmin(sqrt(8.0*t.cs*t.cs*M_MOL*(mi + mj) / (PI*mi*mj)), t.cs);
#endif
    return sqrt(dvr*dvr + dvphi*dvphi + dvz*dvz + vt*vt + vb*vb);
}

// partner-only factors of one slot, held for the whole of an owner's event chain: the owner's position
// factors and the partner's published record do not change during the bath, so these are the values
// vrel_pair_terms would otherwise recompute after every event, by the same operations (the vertical
// overlap of a vertically integrated disk stays per event: holding it too costs the chain kernel a
// wavefront per SIMD)
struct partner_terms
{
    double sj, fj, aj, mj;
};
// the chain holds them where its registers allow two wavefronts per SIMD (the 3D physical kernel; a
// vertically integrated one needs the partner's scale height besides)
constexpr bool HOLD_PARTNER = COAG_KERNEL == 3 && N_Z > 1;

GDX_FN partner_terms make_partner_terms(const owner_terms &t, double size_j)
{
    partner_terms p{0.0, 0.0, 0.0, 0.0};
    if constexpr (COAG_KERNEL == 3)
    {
        p.sj = stokes_partner(t, size_j);
        p.fj = 1.0 / (1.0 + p.sj*p.sj);
        p.aj = p.sj*p.fj;
        p.mj = fmin(p.sj, 0.5);
    }
    else
    {
        (void)t; (void)size_j;
    }
    return p;
}

// vrel_pair_terms from held partner factors: bit for bit the same speed
GDX_FN double vrel_pair_held(const owner_terms &t, const partner_terms &p, double size_j)
{
    double dvr = 2.0*t.vn*(t.ai - p.aj);
    double dvphi = t.vn*(t.fi - p.fj);
    double dvz = t.zo*(t.mi - p.mj);
    double vt = vrel_t_merged(t.ris, t.vg_sq, t.si, p.sj);
    double vb = 0.0;
#ifndef CODE_UNIT
    double mi = grain_mass(t.size_i);
    double mj = grain_mass(size_j);
    vb = fmin(sqrt(8.0*t.cs*t.cs*M_MOL*(mi + mj) / (PI*mi*mj)), t.cs);
#else
    (void)size_j;
#endif
    return sqrt(dvr*dvr + dvphi*dvphi + dvz*dvz + vt*vt + vb*vb);
}

// catastrophic-fragmentation moments (_col_event.cuh:62-73): they depend on the owner size alone
struct cat_moments
{
    double mean, second, maximum;
    GDX_FN void get(double &a, double &b, double &c) const { a = mean; b = second; c = maximum; }
};

// zeros in place of the catastrophic moments: a caller that fills them in afterwards for the pairs that
// took the catastrophic branch (catastrophic_pair), where the terms are rate x moment with the rate
// unchanged (the branch's rate factor is 1)
struct cat_moments_deferred
{
    GDX_FN void get(double &a, double &b, double &c) const { a = 0.0; b = 0.0; c = 0.0; }
};

// the same moments read from shared memory only where a lane needs them (volatile: neither
// forwarded into registers nor hoisted out of the pair loop)
struct cat_moments_lds
{
    const volatile double *p;
    __device__ __forceinline__ void get(double &a, double &b, double &c) const { a = p[0]; b = p[1]; c = p[2]; }
};

GDX_FN cat_moments catastrophic_moments(double si)
{
    cat_moments c{0.0, 0.0, 0.0};
    if constexpr (!CAN_FRAGMENT) return c;
    double L = 0.5*log(div_rn(si, INIT_SMIN, INV_SMIN));
    if (L < 1.0e-03)
    {
        c.mean = L - L*L / 6.0 + L*L*L*L / 360.0;
        c.second = L*L*(4.0/3.0 - L / 3.0 + L*L / 90.0 + L*L*L / 180.0);
    }
    else
    {
        double tail = exp(-L) / (-expm1(-L));
        c.mean = 2.0*(1.0 - L*tail);
        c.second = 8.0 - (4.0*L*L + 8.0*L)*tail;
    }
    c.maximum = 2.0*L;
    return c;
}

// whether outcome_moments_terms takes its catastrophic branch for this partner: the same ratio, the
// same tests in the same order (a NaN ratio falls through to it as there)
GDX_FN bool catastrophic_pair(const owner_terms &t, double sj, bool high_speed)
{
    double ratio = div_rn(sj, t.size_i, t.inv_size);
    double q = ratio*ratio*ratio;
    return high_speed && !(q <= 0.1);
}

// _erosion_outcome_moments with owner-uniform and constant divisors (_col_event.cuh:31-74). The
// sticking (log1p(Gp q)) and erosion remnant (log1p(-Gp q)) forms share one log1p call with the
// argument selected per lane; the catastrophic branch reads the owner's precomputed moments.
template <class CM>
GDX_FN double outcome_moments_terms(const owner_terms &t, const CM &cm, double sj, bool high_speed,
                                    double &mean, double &second, double &maximum)
{
    const double si = t.size_i;
    double ratio = div_rn(sj, si, t.inv_size);
    double q = ratio*ratio*ratio;
    double Gp = sticking_packet(q);
    double lp = ulog1p(high_speed ? -Gp*q : Gp*q);
    if (!high_speed)
    {
        mean = div_rn(lp, 3.0, INV_THREE);
        second = mean*mean;
        maximum = mean;
        return Gp == 1.0 ? 1.0 : 1.0 / Gp;
    }
    if (q <= 0.1)
    {
        double remnant = Gp == 1.0 ? (1.0 - q) : (1.0 - q) / Gp;
        double debris = q;
        double factor = remnant + debris;
        double jr = div_rn(-lp, 3.0, INV_THREE);
        double jd = div_rn(-ulog(q), 3.0, INV_THREE);
        mean = (remnant*jr + debris*jd) / factor;
        second = (remnant*jr*jr + debris*jd*jd) / factor;
        maximum = fmax(jr, jd);
        return factor;
    }
    cm.get(mean, second, maximum);
    return 1.0;
}

// one slot's pair term (_get_col_chain_rate / measure x outcome factor; col_bath_rate.cu:40-57)
struct pair_term
{
    double rate, j1, j2, jmax, vrel, size_j;
};

// partner data of one slot, loaded ahead of its evaluation
struct partner
{
    double size_j, numr_j, hg_j;
    bool valid;
};

GDX_FN partner load_partner(int j, const double2 *res, const double *hgs)
{
    partner q{0.0, 0.0, 0.0, false};
    if (j < 0) return q;
    double2 pj = res[j];
    q.size_j = pj.x;
    q.numr_j = pj.y;
    if constexpr (PARTNER_HG) q.hg_j = hgs[j];
    else (void)hgs;
    q.valid = true;
    return q;
}

template <class CM, bool HELD = false>
GDX_FN pair_term eval_pair_terms(const owner_terms &t, const CM &cm, const partner &q, double lambda_0,
                                 const partner_terms &pt = partner_terms{})
{
    pair_term p{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (!q.valid) return p;
    const double size_i = t.size_i, size_j = q.size_j, numr_j = q.numr_j;
    double vrel = 0.0;
    double rate;
    if constexpr (COAG_KERNEL == 0)
    {
        rate = lambda_0*numr_j;
    }
    else if constexpr (COAG_KERNEL == 1)
    {
        rate = lambda_0*numr_j*(grain_mass(size_i) + grain_mass(size_j));
    }
    else if constexpr (COAG_KERNEL == 2)
    {
        rate = lambda_0*numr_j*grain_mass(size_i)*grain_mass(size_j);
    }
    else
    {
        vrel = HELD ? vrel_pair_held(t, pt, size_j) : vrel_pair_terms(t, size_j);
        rate = numr_j*vrel*PI*(size_i + size_j)*(size_i + size_j) / 4.0;
        if constexpr (N_Z == 1) rate /= sqrt(2.0*PI*(t.hg*t.hg + q.hg_j*q.hg_j));
    }
    double r = div_rn(rate, t.m, t.inv_m);
    double mean = 0.0, second = 0.0, maximum = 0.0;
    r *= outcome_moments_terms(t, cm, size_j, vrel >= V_FRAG, mean, second, maximum);
    p.rate = r;
    p.j1 = r*mean;
    p.j2 = r*second;
    p.jmax = r > 0.0 ? maximum : 0.0;
    p.vrel = vrel;
    p.size_j = size_j;
    return p;
}

// the rate factor of outcome_moments_terms alone (callers that keep no compensators): the same
// operations, without the logarithms of the moments
GDX_FN double outcome_factor(const owner_terms &t, double sj, bool high_speed)
{
    const double si = t.size_i;
    double ratio = div_rn(sj, si, t.inv_size);
    double q = ratio*ratio*ratio;
    double Gp = sticking_packet(q);
    if (!high_speed) return Gp == 1.0 ? 1.0 : 1.0 / Gp;
    if (q <= 0.1)
    {
        double remnant = Gp == 1.0 ? (1.0 - q) : (1.0 - q) / Gp;
        return remnant + q;
    }
    return 1.0;
}

// eval_pair_terms without the jump moments: the same rate bit for bit
template <bool HELD = false>
GDX_FN pair_term eval_pair_rate(const owner_terms &t, const partner &q, double lambda_0,
                                const partner_terms &pt = partner_terms{})
{
    pair_term p{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (!q.valid) return p;
    const double size_i = t.size_i, size_j = q.size_j, numr_j = q.numr_j;
    double vrel = 0.0;
    double rate;
    if constexpr (COAG_KERNEL == 0)
    {
        rate = lambda_0*numr_j;
    }
    else if constexpr (COAG_KERNEL == 1)
    {
        rate = lambda_0*numr_j*(grain_mass(size_i) + grain_mass(size_j));
    }
    else if constexpr (COAG_KERNEL == 2)
    {
        rate = lambda_0*numr_j*grain_mass(size_i)*grain_mass(size_j);
    }
    else
    {
        vrel = HELD ? vrel_pair_held(t, pt, size_j) : vrel_pair_terms(t, size_j);
        rate = numr_j*vrel*PI*(size_i + size_j)*(size_i + size_j) / 4.0;
        if constexpr (N_Z == 1) rate /= sqrt(2.0*PI*(t.hg*t.hg + q.hg_j*q.hg_j));
    }
    double r = div_rn(rate, t.m, t.inv_m);
    r *= outcome_factor(t, size_j, vrel >= V_FRAG);
    p.rate = r;
    p.vrel = vrel;
    p.size_j = size_j;
    return p;
}

// moving logarithmic size bin (inc/swarm/_col_sizes.cuh:12-21)
GDX_FN int size_bin(double size, double lower, double upper)
{
    if (!isfinite(size) || !(size > 0)) return 0;
    double fraction = log(size / lower) / log(upper / lower);
    return max(0, min(COL_BIN_S - 1, static_cast<int>(floor(fraction*COL_BIN_S))));
}

// ---------------------------------------------------------------------------------------------
// wavefront segment helpers: a segment is SEG consecutive lanes holding one owner's K slots
// ---------------------------------------------------------------------------------------------
template <int W>
__device__ __forceinline__ double seg_sum(double v)
{
#pragma unroll
    for (int m = W / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, W);
    return v;
}
template <int W>
__device__ __forceinline__ double seg_max(double v)
{
#pragma unroll
    for (int m = W / 2; m > 0; m >>= 1) v = fmax(v, __shfl_xor(v, m, W));
    return v;
}
template <int W>
__device__ __forceinline__ int seg_last(int v)
{
#pragma unroll
    for (int m = W / 2; m > 0; m >>= 1)
    {
        int o = __shfl_xor(v, m, W);
        v = (v == -2 || o == -2) ? -2 : (v > o ? v : o);
    }
    return v;
}
template <int W>
__device__ __forceinline__ int seg_and(int v)
{
#pragma unroll
    for (int m = W / 2; m > 0; m >>= 1) v = v & __shfl_xor(v, m, W);
    return v;
}

// gfx950 hands workgroups to its 8 XCDs round-robin; each XCD has its own L2. Launching
// XCDS*ceil(n/XCDS) groups and mapping group b to logical block (b % XCDS)*per + b / XCDS gives each
// XCD one contiguous range of logical blocks, so owners that share partners share an L2.
constexpr int XCDS = 8;
__host__ __device__ constexpr int xcd_grid(int logical_blocks)
{
    return XCDS*((logical_blocks + XCDS - 1) / XCDS);
}
__device__ __forceinline__ int xcd_block()
{
    int per = int(gridDim.x) / XCDS;
    int b = int(blockIdx.x);
    return (b % XCDS)*per + b / XCDS;
}

__device__ __forceinline__ unsigned long long dbits(double v) { return static_cast<unsigned long long>(__double_as_longlong(v)); }
__device__ __forceinline__ double bitsd(unsigned long long v) { return __longlong_as_double(static_cast<long long>(v)); }

}  // namespace gdx

#endif  // GDX_DEVICE_H
