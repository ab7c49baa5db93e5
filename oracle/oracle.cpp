// CPU oracle for the swarm collision operator -- TEST INFRASTRUCTURE ONLY.
//
// Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
// Built per model directory with g++ -ffp-contract=off so every float and double expression is
// evaluated in source order without fused multiply-adds. See oracle.h for what pins each part.
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "const_defs.h"

using namespace gdx;

namespace {

// Planted mistakes for the pin tests (tests/test_oracle_pins.py builds the oracle with -DORC_PLANT=n
// and checks that the pins reject every variant); the ordinary build defines no ORC_PLANT.
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif
constexpr int PLANT = ORC_PLANT;

constexpr double PI = 3.14159265358979323846;
constexpr int DIM = (N_Z > 1) ? 3 : 2;
constexpr bool X_WEDGE = N_X > 1
    && static_cast<float>(X_MAX) - static_cast<float>(X_MIN) < 6.28318530717958647692f - 1.0e-06f;
constexpr int LOCAL_GROUPS = ((N_X > 1) ? COL_BIN_X : 1)*COL_BIN_Y*((N_Z > 1) ? COL_BIN_Z : 1);
constexpr int LOCAL_WORDS = (LOCAL_GROUPS + 31) / 32;
constexpr int RAW_BINS = LOCAL_GROUPS*COL_BIN_S;

// ============================================================================================
// random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), written from the paper
// ============================================================================================
inline void philox_round(uint32_t c[4], const uint32_t k[2])
{
    uint64_t p0 = uint64_t(0xD2511F53u)*c[0];
    uint64_t p1 = uint64_t(0xCD9E8D57u)*c[2];
    uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

inline void philox10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r)
    {
        if (r) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
        philox_round(c, k);
    }
    for (int i = 0; i < 4; ++i) out[i] = c[i];
}

// one uniform deviate in (0,1) for draw number `counter` of particle `pidx` under `seed`:
// counter words {counter lo, counter hi, seed hi, tag}, key {pidx, seed lo}; 52 random bits
inline double uniform(uint32_t pidx, uint64_t seed, uint64_t counter)
{
    uint32_t ctr[4] = {uint32_t(counter), uint32_t(counter >> 32), uint32_t(seed >> 32), 0x47445843u};
    uint32_t key[2] = {pidx, uint32_t(seed)};
    uint32_t out[4];
    philox10(ctr, key, out);
    uint64_t m = (uint64_t(out[1]) << 32) | out[0];
    return (double(m >> 12) + 0.5)*0x1p-52;
}

// _get_col_uniform (inc/swarm/_col_rates.cuh:145-149): strictly below one
inline double col_uniform(uint32_t pidx, uint64_t seed, uint64_t &counter)
{
    double u = uniform(pidx, seed, counter);
    counter++;
    return std::fmin(u, std::nextafter(1.0, 0.0));
}

// ============================================================================================
// disk and grain profiles (inc/swarm/param_phys.cuh, param_grid.cuh)
// ============================================================================================
inline double grain_mass(double size) { return PI*RHO_0*size*size*size / 6.0; }         // param_phys.cuh:16-18
inline double omegaK(double R) { return std::sqrt(G*M_S / R / R / R); }                  // param_phys.cuh:66-68
inline double hg(double R) { return ASPR_0*std::pow(R / R_0, 0.5*(IDX_Q + 1.0)); }       // param_phys.cuh:70-72
inline double eta(double R, double Z, double h)                                          // param_phys.cuh:75-77
{ return -0.5*((IDX_P + 0.5*IDX_Q - 1.5)*h*h + (PLANT == 7 ? -IDX_Q : IDX_Q)*(1.0 - R / std::sqrt(R*R + Z*Z))); }
inline double gas_strat(double R, double Z, double h)                                    // param_phys.cuh:80-82
{ return std::exp((R / std::sqrt(R*R + Z*Z) - 1.0) / (h*h)); }
inline double sigma_g(double R) { return SIGMA_0*std::pow(R / R_0, IDX_P); }             // param_phys.cuh:86-87
inline double cs(double R, double h) { return h*omegaK(R)*R; }                            // param_phys.cuh:95-97
inline double alpha_of(double R, double h)                                               // param_phys.cuh:113-122
{
#ifdef CONST_NU
    return NU / (h*h*R*R*omegaK(R));
#else
    (void)R; (void)h;
    return ALPHA;
#endif
}
inline double stokes(double R, double Z, double h, double size)                          // param_phys.cuh:163-210
{
    double st = STOKES_0*(size / S_0);
#ifndef CONST_ST
    if (PLANT == 4) st *= std::pow(R / R_0, IDX_P);
    else st /= std::pow(R / R_0, IDX_P);
    st /= gas_strat(R, Z, h);
#else
    (void)R; (void)Z; (void)h;
#endif
    return st;
}
inline double cyl_R(double y, double z) { return (N_Z > 1) ? y*std::sin(z) : y; }       // param_grid.cuh:23
inline double cyl_Z(double y, double z) { return (N_Z > 1) ? y*std::cos(z) : 0.0; }     // param_grid.cuh:27

inline double re_inv_sqrt(double R, double alpha, double sig)                            // _collision.cuh:136-150
{
    (void)R;
    double reynolds = 1.0;
#ifdef CODE_UNIT
    double alpha_0 = alpha_of(R_0, ASPR_0);
    reynolds = REYNOLDS_0*(alpha / alpha_0)*(sig / SIGMA_0);
#else
    reynolds = 0.5*alpha*sig*X_SEC / M_MOL;
#endif
    return 1.0 / std::sqrt(reynolds);
}

// Ormel & Cuzzi (2007) regimes, _collision.cuh:152-273
double vrel_t(double R, double stokes_i, double stokes_j, double h_g, double sig)
{
    double c_s = cs(R, h_g);
    double alpha = alpha_of(R, h_g);
    double ris = re_inv_sqrt(R, alpha, sig);
    double vg_sq = 1.5*alpha*c_s*c_s;
    double sl, ss;
    if (stokes_i >= stokes_j) { sl = stokes_i; ss = stokes_j; }
    else { sl = stokes_j; ss = stokes_i; }
    double eps = ss / sl;
    double y_a = 1.6;
    double y_s = 1.6015125;
    y_s += -0.63119577*sl;
    y_s +=  0.32938936*sl*sl;
    y_s += -0.29847604*sl*sl*sl;
    double v = 0.0;
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
        double coeff = 0.0;
        coeff  = (sl - ss) / (sl + ss);
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
    if (v < 0.0) throw std::runtime_error("negative vrel_sq in vrel_t");
    return std::sqrt(v);
}

#ifndef CODE_UNIT
double vrel_b(double R, double size_i, double size_j, double h)                          // _collision.cuh:117-133
{
    double c_s = cs(R, h);
    double m_i = grain_mass(size_i);
    double m_j = grain_mass(size_j);
    double v = std::sqrt(8.0*c_s*c_s*M_MOL*(m_i + m_j) / (PI*m_i*m_j));
    return std::min(v, c_s);
}
#endif

// query-local pair speed at the owner's position, _collision.cuh:275-325
double vrel_pair(double y, double z, double size_i, double size_j)
{
    double R = cyl_R(y, z);
    double Z = cyl_Z(y, z);
    double h = hg(R);
    double omega = omegaK(R);
    double si = stokes(R, Z, h, size_i);
    double sj = stokes(R, Z, h, size_j);
    double vn = -eta(R, Z, h)*R*omega;
    double fi = 1.0 / (1.0 + si*si);
    double fj = 1.0 / (1.0 + sj*sj);
    double dvr = (PLANT == 1 ? 1.0 : 2.0)*vn*(si*fi - sj*fj);
    double dvphi = vn*(fi - fj);
    double cap = PLANT == 2 ? 1.0 : 0.5;
    double dvz = Z*omega*(std::fmin(si, cap) - std::fmin(sj, cap));
    double sigma_local = sigma_g(R)*gas_strat(R, Z, h);
    double vt = vrel_t(R, si, sj, h, sigma_local);
    double vb = 0.0;
#ifndef CODE_UNIT
    vb = vrel_b(R, size_i, size_j, h);
#endif
    return std::sqrt(dvr*dvr + dvphi*dvphi + dvz*dvz + vt*vt + vb*vb);
}

// KNN ball measure clipped by radial and polar domain faces, _collision.cuh:66-115
double ball_measure(double y, double z, double radius)
{
#ifdef GDX_UNIT_MEASURE
    (void)y; (void)z;
    return radius > 0.0 ? 1.0 : 0.0;
#else
    if (radius <= 0.0) return 0.0;
    if (N_X == 1 && N_Z == 1)
    {
        double R_lo = std::fmax(Y_MIN, y - radius);
        double R_hi = std::fmin(Y_MAX, y + radius);
        return (R_hi > R_lo) ? PI*(R_hi*R_hi - R_lo*R_lo) : 0.0;
    }
    int dim = 1 + int(N_X > 1) + int(N_Z > 1);
    double measure = (dim == 1) ? 2.0*radius : (dim == 2) ? PI*radius*radius
        : 4.0*PI*radius*radius*radius / 3.0;
    double distances[4] = {y - Y_MIN, Y_MAX - y, 1.0e+100, 1.0e+100};
    if (N_Z > 1)
    {
        distances[2] = y*(z - Z_MIN);
        distances[3] = y*(Z_MAX - z);
    }
    int bounds = (N_Z > 1) ? 4 : 2;
    for (int b = 0; b < bounds; b++)
    {
        double d = std::fmax(0.0, distances[b]);
        if (d >= radius) continue;
        if (dim == 1)
        {
            measure *= 0.5*(1.0 + d / radius);
        }
        else if (dim == 2)
        {
            double cap = radius*radius*std::acos(d / radius) - d*std::sqrt(radius*radius - d*d);
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

// outcome packets and jump moments, _col_event.cuh:15-74
inline double sticking_packet(double q, bool fragmentation)
{ return !fragmentation && q > 0.0 && q <= 1.0e-06 ? std::fmax(1.0, std::floor(1.0e-04 / q)) : 1.0; }
inline double sticking_mass_ratio(double size_i, double size_j)
{ double ratio = size_j / size_i; return ratio*ratio*ratio; }

double outcome_moments(double si, double sj, bool high_speed, double &mean, double &second, double &maximum)
{
    double q = sticking_mass_ratio(si, sj);
    double G = sticking_packet(q, false);
    if (!high_speed)
    {
        mean = std::log1p(G*q) / 3.0;
        second = mean*mean;
        maximum = mean;
        return 1.0 / G;
    }
    if (q <= 0.1)
    {
        double remnant = (1.0 - q) / G;
        double debris = q;
        double factor = remnant + debris;
        double jr = -std::log1p(-G*q) / 3.0;
        double jd = -std::log(q) / 3.0;
        mean = (remnant*jr + debris*jd) / factor;
        second = (remnant*jr*jr + debris*jd*jd) / factor;
        maximum = std::fmax(jr, jd);
        return factor;
    }
    double L = 0.5*std::log(si / INIT_SMIN);
    if (L < 1.0e-03)
    {
        mean = L - L*L / 6.0 + L*L*L*L / 360.0;
        second = L*L*(4.0/3.0 - L / 3.0 + L*L / 90.0 + L*L*L / 180.0);
    }
    else
    {
        double tail = std::exp(-L) / (-std::expm1(-L));
        mean = 2.0*(1.0 - L*tail);
        second = 8.0 - (4.0*L*L + 8.0*L)*tail;
    }
    maximum = 2.0*L;
    return 1.0;
}

// _col_event.cuh:79-109
double sample_outcome(double si, double sj, bool high_speed, double u, int &category, double &log_mass)
{
    double q = sticking_mass_ratio(si, sj);
    double G = sticking_packet(q, false);
    if (!high_speed)
    {
        category = q <= 1.0e-06 ? 0 : q <= 1.0e-04 ? 1 : q <= 1.0e-02 ? 2 : 3;
        log_mass = std::log1p(G*q);
        return std::cbrt(si*si*si + G*sj*sj*sj);
    }
    if (q <= 0.1)
    {
        double factor = (1.0 - q) / G + q;
        if (u < q / factor)
        {
            category = 6;
            log_mass = std::log(q);
            return sj;
        }
        category = 5;
        log_mass = std::log1p(-G*q);
        return si*std::cbrt(1.0 - G*q);
    }
    category = 4;
    double lower = std::sqrt(INIT_SMIN);
    double root = lower + u*(std::sqrt(si) - lower);
    double size = std::fmin(si, std::fmax(INIT_SMIN, root*root));
    log_mass = 3.0*std::log(size / si);
    return size;
}

// _get_col_chain_rate (inc/swarm/_col_rates.cuh:167-223), owner i at (y_i, z_i)
double chain_rate(double y_i, double z_i, double y_j, double z_j, double size_i, double size_j,
                  double numr_j, double lambda_0, double &vrel)
{
    vrel = 0.0;
    if (COAG_KERNEL == 0)
    {
        return lambda_0*numr_j;
    }
    else if (COAG_KERNEL == 1)
    {
        return lambda_0*numr_j*(grain_mass(size_i) + grain_mass(size_j));
    }
    else if (COAG_KERNEL == 2)
    {
        return lambda_0*numr_j*grain_mass(size_i)*grain_mass(size_j);
    }
    vrel = vrel_pair(y_i, z_i, size_i, size_j);
    double rate = numr_j*vrel*PI*(size_i + size_j)*(size_i + size_j) / 4.0;
    if (N_Z == 1)
    {
        double R_i = cyl_R(y_i, z_i);
        double R_j = cyl_R(y_j, z_j);
        double H_gi = R_i*hg(R_i);
        double H_gj = PLANT == 3 ? R_i*hg(R_i) : R_j*hg(R_j);
        rate /= std::sqrt(2.0*PI*(H_gi*H_gi + H_gj*H_gj));
    }
    return rate;
}

// ============================================================================================
// geometry
// ============================================================================================
inline bool is_active(double y, double z)                                                 // _transport.cuh:37-49
{
    if (y < Y_MIN || y >= Y_MAX) return false;
#ifdef HALF_DISK
    if (N_Z > 1 && (z < Z_MIN || z > Z_MAX)) return false;
#else
    if (N_Z > 1 && (z < Z_MIN || z >= Z_MAX)) return false;
#endif
    return true;
}

// _expand_morton_3d / _get_morton_key (inc/swarm/morton/morton_types.cuh:46-65)
inline uint64_t expand3(uint32_t v)
{
    uint64_t b = v & 0x1fffffu;
    b = (b | b << 32) & 0x001f00000000ffffULL;
    b = (b | b << 16) & 0x001f0000ff0000ffULL;
    b = (b | b << 8)  & 0x100f00f00f00f00fULL;
    b = (b | b << 4)  & 0x10c30c30c30c30c3ULL;
    b = (b | b << 2)  & 0x1249249249249249ULL;
    return b;
}

inline float point_d2(const float *a, const float *b)                                    // morton_types.cuh:75-82
{
    float dx = a[0] - b[0];
    float dy = a[1] - b[1];
    float dz = a[2] - b[2];
    return dx*dx + dy*dy + dz*dz;
}

inline bool nb_less(float da, int ia, float db, int ib)                                   // morton_types.cuh:68-72
{ return da < db || (da == db && ia < ib); }

// morton_node padded lower distance (inc/swarm/morton/morton_index.cuh:296-314)
inline float node_d2(const float *q, const float *lower, float width, int dim, int max_level)
{
    float ux = lower[0] + width, uy = lower[1] + width, uz = lower[2] + width;
    float scale = std::fmax(1.0f, std::fmax(std::fmax(std::fabs(lower[0]), std::fabs(ux)),
        std::fmax(std::fmax(std::fabs(lower[1]), std::fabs(uy)), std::fmax(std::fabs(lower[2]), std::fabs(uz)))));
    float pad = 2.0f*float(max_level + 2)*FLT_EPSILON*scale;
    float dx = std::fmax(std::fmax(lower[0] - q[0] - pad, 0.0f), q[0] - ux - pad);
    float dy = std::fmax(std::fmax(lower[1] - q[1] - pad, 0.0f), q[1] - uy - pad);
    float dz = (dim == 2) ? 0.0f : std::fmax(std::fmax(lower[2] - q[2] - pad, 0.0f), q[2] - uz - pad);
    return dx*dx + dy*dy + dz*dz;
}

int child_end(const uint64_t *keys, int begin, int end, int shift, int code)             // morton_index.cuh:63-73
{
    while (begin < end)
    {
        int mid = begin + (end - begin) / 2;
        if (int((keys[mid] >> shift) & 7ULL) < code) begin = mid + 1;
        else end = mid;
    }
    return begin;
}

int space_bin(double x, double y, double z)                                                // col_space_bin.cu:11-35
{
    int bin_x = 0;
    int bin_y = int(std::floor((y - Y_MIN) / (Y_MAX - Y_MIN)*COL_BIN_Y));
    int bin_z = 0;
    if (N_X > 1) bin_x = int(std::floor((x - X_MIN) / (X_MAX - X_MIN)*COL_BIN_X));
    if (N_Z > 1)
    {
        double extent_z = Z_MAX - Z_MIN;
        bin_z = int(std::floor((z - Z_MIN) / extent_z*COL_BIN_Z));
    }
    bin_x = (bin_x < 0) ? 0 : ((bin_x >= COL_BIN_X) ? COL_BIN_X - 1 : bin_x);
    bin_y = (bin_y < 0) ? 0 : ((bin_y >= COL_BIN_Y) ? COL_BIN_Y - 1 : bin_y);
    bin_z = (bin_z < 0) ? 0 : ((bin_z >= COL_BIN_Z) ? COL_BIN_Z - 1 : bin_z);
    int count_x = (N_X > 1) ? COL_BIN_X : 1;
    return bin_x + count_x*(bin_y + COL_BIN_Y*bin_z);
}

// ============================================================================================
// host controller (inc/swarm/_col_bound.cuh) and scheduler (inc/swarm/_col_sched.cuh)
// ============================================================================================
struct change_bound { double duration; int reason; };

change_bound change_limit(double A, double B, double epsilon, double horizon)             // _col_bound.cuh:22-33
{
    if (!std::isfinite(A) || A < 0 || !std::isfinite(B) || B < 0
        || !std::isfinite(epsilon) || !(epsilon > 0) || !std::isfinite(horizon) || !(horizon > 0))
        throw std::runtime_error("invalid change-based refresh inputs");
    change_bound result{horizon, 0};
    if (A > 0 && epsilon / A < result.duration) result = {epsilon / A, 1};
    if (B > 0 && epsilon*epsilon / B < result.duration) result = {epsilon*epsilon / B, 2};
    if (!(result.duration > 0)) throw std::runtime_error("change-based timestep underflow");
    return result;
}

int build_binmap(const std::vector<int> &raw_count, std::vector<int> &raw_to_merged)      // _col_bound.cuh:45-98
{
    int spatial_count = int(raw_count.size()) / COL_BIN_S;
    raw_to_merged.assign(raw_count.size(), -1);
    int merged_count = 0;
    for (int s = 0; s < spatial_count; s++)
    {
        int last_occupied = -1;
        for (int b = 0; b < COL_BIN_S; b++) if (raw_count[s*COL_BIN_S + b] > 0) last_occupied = b;
        if (last_occupied < 0)
        {
            int m = merged_count++;
            for (int b = 0; b < COL_BIN_S; b++) raw_to_merged[s*COL_BIN_S + b] = m;
            continue;
        }
        int begin = 0, previous = -1;
        while (begin <= last_occupied)
        {
            int end = begin, count = 0;
            while (end <= last_occupied && count < COL_BIN_MIN) count += raw_count[s*COL_BIN_S + end++];
            bool merge_tail = end > last_occupied && count < COL_BIN_MIN && previous >= 0;
            int m = merge_tail ? previous : merged_count++;
            for (int b = begin; b < end; b++) raw_to_merged[s*COL_BIN_S + b] = m;
            previous = m;
            begin = end;
        }
        if (last_occupied + 1 < COL_BIN_S)
        {
            int m = merged_count++;
            for (int b = last_occupied + 1; b < COL_BIN_S; b++) raw_to_merged[s*COL_BIN_S + b] = m;
        }
    }
    return merged_count;
}

struct rate_bin { double mass = 0, weighted_rate = 0, weighted_change = 0, weighted_second = 0; int owner_count = 0, invalid_count = 0; };
struct audit_bin
{
    double mass = 0, predicted_f = 0, predicted_e = 0, predicted_var_f = 0, predicted_var_e = 0,
        predicted_g = 0, predicted_var_g = 0, maximum_weight = 0, maximum_g_jump = 0, touched = 0,
        event_weight = 0, growth = 0, end_mass = 0;
    int owner_count = 0, invalid_count = 0;
};
struct bath_state { double limit_scale = 1.0; int activity_streak = 0, quiet_streak = 0, floor_streak = 0; };
struct bath_result
{
    double max_f = 0, max_e = 0, max_touched = 0, max_events = 0, max_g = 0, max_g_upper = 0, d_bath = 0;
    bool activity_overshoot = false, distribution_overshoot = false, persistent_overshoot = false;
};

double choose_bath(const std::vector<rate_bin> &bin, double remaining, double limit_scale, int *binding) // _col_bound.cuh:101-123
{
    double duration = std::min(remaining, COL_BATH_MAX);
    double tolerance = COL_BATH_EPS*limit_scale;
    int reason = 0;
    for (const auto &v : bin)
    {
        if (v.invalid_count) throw std::runtime_error("invalid change-based rate moments");
        if (!(v.mass > 0)) continue;
        auto bound = change_limit(v.weighted_change / v.mass, v.weighted_second / v.mass, tolerance, duration);
        if (bound.duration < duration) { duration = bound.duration; reason = bound.reason; }
    }
    if (binding) *binding = reason;
    return duration;
}

bath_result finish_bath(const std::vector<audit_bin> &bin, bath_state &state)           // _col_bound.cuh:126-216
{
    bath_result result;
    int active_count = 0;
    double total_mass = 0.0;
    for (const auto &v : bin)
    {
        if (v.invalid_count != 0) throw std::runtime_error("collision bath controller returned an invalid audit state");
        if (v.mass > 0.0) active_count++;
        total_mass += v.mass;
    }
    if (!(total_mass > 0.0)) throw std::runtime_error("collision bath controller found no active represented mass");
    double L = std::log(2.0*double(active_count) / COL_BATH_ALPHA);
    double mass_difference = 0.0;
    for (const auto &v : bin)
    {
        mass_difference += std::abs(v.end_mass - v.mass);
        if (!(v.mass > 0.0)) continue;
        double mass_sq = v.mass*v.mass;
        double pf = v.predicted_f / v.mass, pe = v.predicted_e / v.mass, pg = v.predicted_g / v.mass;
        double touched = v.touched / v.mass, events = v.event_weight / v.mass, growth = v.growth / v.mass;
        double wmax = v.maximum_weight / v.mass, jmax = v.maximum_g_jump / v.mass;
        double upper_f = std::min(1.0, pf + std::sqrt(2.0*v.predicted_var_f / mass_sq*L) + wmax*L / 3.0);
        double upper_e = pe + std::sqrt(2.0*v.predicted_var_e / mass_sq*L) + wmax*L / 3.0;
        double upper_g = pg + std::sqrt(2.0*v.predicted_var_g / mass_sq*L) + jmax*L / 3.0;
        result.max_f = std::max(result.max_f, pf);
        result.max_e = std::max(result.max_e, pe);
        result.max_touched = std::max(result.max_touched, touched);
        result.max_events = std::max(result.max_events, events);
        result.max_g = std::max(result.max_g, growth);
        result.max_g_upper = std::max(result.max_g_upper, upper_g);
        result.activity_overshoot = result.activity_overshoot || std::min(touched, 1.0) > upper_f || events > upper_e;
        result.distribution_overshoot = result.distribution_overshoot || growth > std::max(COL_BATH_EPS, upper_g);
    }
    result.d_bath = mass_difference / (2.0*total_mass);
    result.distribution_overshoot = result.distribution_overshoot || result.d_bath > COL_BATH_EPS;
    bool issue = result.activity_overshoot || result.distribution_overshoot;
    bool at_floor = state.limit_scale <= 0.25*(1.0 + 8.0*std::numeric_limits<double>::epsilon());
    if (issue)
    {
        state.activity_streak++;
        state.quiet_streak = 0;
        state.floor_streak = at_floor ? state.floor_streak + 1 : 0;
    }
    else
    {
        state.activity_streak = 0;
        state.quiet_streak++;
        state.floor_streak = 0;
    }
    if (result.distribution_overshoot || state.activity_streak >= 2)
    {
        state.limit_scale = std::max(0.25, 0.5*state.limit_scale);
        state.activity_streak = 0;
    }
    else if (state.quiet_streak >= 3)
    {
        state.limit_scale = std::min(1.0, 1.25*state.limit_scale);
        state.quiet_streak = 0;
    }
    result.persistent_overshoot = state.floor_streak >= 2;
    return result;
}

struct schedule                                                                              // _col_sched.cuh:17-183
{
    double horizon;
    std::vector<std::pair<int, int>> links;
    std::vector<int> level;
    std::vector<uint64_t> step, next;
    uint64_t end;

    schedule(double h, const std::vector<double> &requested, const std::vector<uint32_t> &edges)
        : horizon(h), level(requested.size(), 0), step(requested.size()), next(requested.size(), 0)
    {
        if (!(h > 0) || !std::isfinite(h)) throw std::runtime_error("invalid local collision horizon");
        const int n = int(level.size());
        const int words = (n + 31) / 32;
        for (int c = 0; c < n; ++c)
        {
            if (!(requested[c] > 0) || !std::isfinite(requested[c])) throw std::runtime_error("invalid local collision timestep");
            double dt = h;
            while (dt > requested[c])
            {
                if (++level[c] > 52) throw std::runtime_error("local collision timestep exceeds time resolution");
                dt *= 0.5;
            }
        }
        for (int c = 0; c < n; ++c)
            for (int d = c + 1; d < n; ++d)
                if ((edges[c*words + d / 32] & (1u << (d % 32))) || (edges[d*words + c / 32] & (1u << (c % 32))))
                    links.emplace_back(c, d);
        bool changed;
        do
        {
            changed = false;
            for (int c = 0; c < n; ++c)
                for (int d = 0; d < n; ++d)
                {
                    if (!(edges[c*words + d / 32] & (1u << (d % 32)))) continue;
                    if (level[c] < level[d] - 4) { level[c] = level[d] - 4; changed = true; }
                    if (level[d] < level[c] - 4) { level[d] = level[c] - 4; changed = true; }
                }
        } while (changed);
        end = uint64_t(1) << 52;
        for (int c = 0; c < n; ++c) step[c] = uint64_t(1) << (52 - level[c]);
    }
    uint64_t time() const { return *std::min_element(next.begin(), next.end()); }
    double seconds(uint64_t tick) const { return horizon*(double(tick) / double(end)); }
    std::vector<int> due(uint64_t tick) const
    {
        std::vector<int> out;
        for (int c = 0; c < int(next.size()); ++c) if (next[c] == tick) out.push_back(c);
        return out;
    }
    void adapt(uint64_t tick, const std::vector<int> &groups, const std::vector<double> &requested, const std::vector<bool> &passed)
    {
        const int n = int(level.size());
        std::vector<bool> active(n, false);
        for (int c : groups) active[c] = true;
        std::vector<int> cap(n, 52), candidate = level;
        for (int c = 0; c < n; ++c) if (!active[c]) cap[c] = level[c];
        bool changed;
        do
        {
            changed = false;
            for (const auto &e : links)
            {
                int c = e.first, d = e.second;
                if (active[c] && cap[c] > cap[d] + 4) { cap[c] = cap[d] + 4; changed = true; }
                if (active[d] && cap[d] > cap[c] + 4) { cap[d] = cap[c] + 4; changed = true; }
            }
        } while (changed);
        for (int c : groups)
        {
            if (!(requested[c] > 0) || !std::isfinite(requested[c])) throw std::runtime_error("invalid adaptive collision timestep");
            int want = 0;
            double h = horizon;
            while (h > requested[c])
            {
                if (++want > 52) throw std::runtime_error("adaptive collision time resolution exceeded");
                h *= 0.5;
            }
            candidate[c] = passed[c] ? want : std::max(want, level[c]);
            while (tick % (uint64_t(1) << (52 - candidate[c]))) ++candidate[c];
            candidate[c] = std::min(candidate[c], cap[c]);
        }
        do
        {
            changed = false;
            for (const auto &e : links)
            {
                int c = e.first, d = e.second;
                if (active[c] && candidate[c] < candidate[d] - 4) { candidate[c] = candidate[d] - 4; changed = true; }
                if (active[d] && candidate[d] < candidate[c] - 4) { candidate[d] = candidate[c] - 4; changed = true; }
            }
        } while (changed);
        for (int c : groups)
        {
            level[c] = candidate[c];
            step[c] = uint64_t(1) << (52 - level[c]);
        }
    }
    void advance(const std::vector<int> &groups) { for (int c : groups) next[c] += step[c]; }
};

// persistent controller memory: bath states (_col_chain.cuh:46) and moving size bounds (col_size_zero.cu:8-9)
struct ctl_t
{
    bath_state state[LOCAL_GROUPS];
    double lower[LOCAL_GROUPS];
    double upper[LOCAL_GROUPS];
};

inline int size_bin(const ctl_t &ctl, double size, int group)                              // _col_sizes.cuh:12-21
{
    if (!std::isfinite(size) || !(size > 0)) return 0;
    double fraction = std::log(size / ctl.lower[group]) / std::log(ctl.upper[group] / ctl.lower[group]);
    return std::max(0, std::min(COL_BIN_S - 1, int(std::floor(fraction*COL_BIN_S))));
}

// K-slot reduction in the reference's generic fold order (col_bath_rate.cu:97-118)
template <typename T, typename F>
void fold(std::vector<T> &w, F combine)
{
    for (int count = N_K; count > 1; count = (count + 1) / 2)
    {
        int half = (count + 1) / 2;
        for (int j = 0; j < count / 2; j++) w[j] = combine(w[j], w[j + half]);
    }
}

struct pair_work { std::vector<double> rate, j1, j2, jmax, vrel; std::vector<int> last; std::vector<unsigned char> valid; };

// one owner's K pair terms against the frozen reservoir (col_bath_rate.cu:30-58, col_chain_run.cu:136-165)
void owner_pairs(int i, double size_i, const double *y, const double *z, const int *nbr,
                 const double *measure, const double *size_old, const double *numr_old, double lambda_0,
                 bool bath_start_filter, const uint8_t *active, pair_work &w)
{
    for (int k = 0; k < N_K; k++)
    {
        w.rate[k] = w.j1[k] = w.j2[k] = w.jmax[k] = w.vrel[k] = 0.0;
        int code = nbr[size_t(i)*N_K + k];
        if (code < 0) continue;
        if (bath_start_filter && (active[i] == 0 || !(measure[i] > 0.0))) continue;
        int j = code / 3;
        double size_j = size_old[j];
        double vrel = 0.0;
        double r = chain_rate(y[i], z[i], y[j], z[j], size_i, size_j, numr_old[j], lambda_0, vrel) / measure[i];
        double mean = 0, second = 0, maximum = 0;
        r *= outcome_moments(size_i, size_j, vrel >= V_FRAG, mean, second, maximum);
        w.rate[k] = r;
        w.j1[k] = r*mean;
        w.j2[k] = r*second;
        w.jmax[k] = r > 0.0 ? maximum : 0.0;
        w.vrel[k] = vrel;
    }
}

struct owner_totals { double rate, j1, j2, jmax; int last; bool valid; };

owner_totals reduce_pairs(pair_work &w, std::vector<double> &tmp)
{
    owner_totals t;
    std::vector<int> last(N_K);
    std::vector<unsigned char> valid(N_K);
    for (int k = 0; k < N_K; k++)
    {
        valid[k] = std::isfinite(w.rate[k]) && w.rate[k] >= 0.0;
        last[k] = (!std::isfinite(w.rate[k]) || w.rate[k] < 0.0) ? -2 : (w.rate[k] > 0.0 ? k : -1);
    }
    tmp = w.rate; fold(tmp, [](double a, double b) { return a + b; }); t.rate = tmp[0];
    tmp = w.j1; fold(tmp, [](double a, double b) { return a + b; }); t.j1 = tmp[0];
    tmp = w.j2; fold(tmp, [](double a, double b) { return a + b; }); t.j2 = tmp[0];
    tmp = w.jmax; fold(tmp, [](double a, double b) { return std::fmax(a, b); }); t.jmax = tmp[0];
    fold(last, [](int a, int b) { return (a == -2 || b == -2) ? -2 : (a > b ? a : b); }); t.last = last[0];
    fold(valid, [](unsigned char a, unsigned char b) { return (unsigned char)(a && b); }); t.valid = valid[0];
    return t;
}

// partner of an event (col_chain_run.cu:299-320): the first slot in slot order whose running sum of
// positive pair rates exceeds target = U * total; on roundoff (no slot passes) the last positive slot
int partner_walk(const double *rate, int k_count, double target, int last_positive)
{
    double cumulative = 0.0;
    for (int k = 0; k < k_count; k++)
    {
        if (rate[k] <= 0.0) continue;
        cumulative += rate[k];
        if (PLANT == 5 ? cumulative >= target : cumulative > target) return k;
    }
    if (PLANT == 6)
        for (int k = 0; k < k_count; k++)
            if (rate[k] > 0.0) return k;
    return last_positive;
}

struct owner_bath
{
    double time = 0, hazard = 0, j1 = 0, j2 = 0, jmax = 0;
    int events = 0, error = 0;
    bool complete = false, chained = false;
    long long cat_count[7] = {0, 0, 0, 0, 0, 0, 0};
    double cat_log[7] = {0, 0, 0, 0, 0, 0, 0};
};

// col_skip_scan (col_skip_scan.cu:13-59) followed by col_chain_run (col_chain_run.cu:21-404) run to
// completion: the reference's continuation cap only splits the same sequence across launches
void run_owner(int i, double end, const double *y, const double *z, double *size, double *numr,
               const uint8_t *active, const int *nbr, const double *measure, const double *size_old,
               const double *numr_old, double lambda_0, const double cached[4], uint64_t *rng, uint64_t seed,
               owner_bath &b, pair_work &w, std::vector<double> &tmp, long long &pair_evals)
{
    if (!active[i] || !(measure[i] > 0.0)) { b.time = end; b.complete = true; return; }
    uint32_t pidx = uint32_t(i);
    double size_i = size[i], numr_i = numr[i];
    uint64_t ctr = rng[i];
    bool keep_running = true;
    if (b.time == 0.0 && b.events == 0)
    {
        double c_rate = cached[0];
        if (std::isfinite(c_rate) && c_rate >= 0.0)
        {
            if (c_rate == 0.0) { b.time = end; b.complete = true; return; }
            uint64_t trial = ctr;
            double wait = -std::log(col_uniform(pidx, seed, trial)) / c_rate;
            double remaining = end - b.time;
            if (std::isfinite(wait) && wait > 0.0 && wait >= remaining)
            {
                b.hazard += c_rate*remaining;
                b.j1 += cached[1]*remaining;
                b.j2 += cached[2]*remaining;
                b.jmax = std::fmax(b.jmax, cached[3]);
                b.time = end;
                b.complete = true;
                rng[i] = trial;
                return;
            }
        }
    }
    b.chained = true;
    while (keep_running)
    {
        owner_pairs(i, size_i, y, z, nbr, measure, size_old, numr_old, lambda_0, false, active, w);
        pair_evals += N_K;
        owner_totals t = reduce_pairs(w, tmp);
        if (t.last == -2) { b.error = 1; break; }
        if (!std::isfinite(t.rate) || t.rate < 0.0) { b.error = 2; break; }
        if (t.rate == 0.0) { b.time = end; break; }
        double wait = -std::log(col_uniform(pidx, seed, ctr)) / t.rate;
        double remaining = end - b.time;
        if (!std::isfinite(wait) || !(wait > 0.0)) { b.error = 3; break; }
        if (wait >= remaining)
        {
            b.hazard += t.rate*remaining;
            b.j1 += t.j1*remaining;
            b.j2 += t.j2*remaining;
            b.jmax = std::fmax(b.jmax, t.jmax);
            b.time = end;
            break;
        }
        double event_time = b.time + wait;
        if (!(event_time > b.time)) { b.error = 4; break; }
        b.hazard += t.rate*wait;
        b.j1 += t.j1*wait;
        b.j2 += t.j2*wait;
        b.jmax = std::fmax(b.jmax, t.jmax);
        double target = col_uniform(pidx, seed, ctr)*t.rate;
        int slot = partner_walk(w.rate.data(), N_K, target, t.last);
        if (slot < 0) { b.error = 5; break; }
        int j = nbr[size_t(i)*N_K + slot] / 3;
        double size_j = size_old[j];
        double vrel = 0.0;
        if (COAG_KERNEL == 3) vrel = vrel_pair(y[i], z[i], size_i, size_j);
        double mass_before = numr_i*size_i*size_i*size_i;
        bool high_speed = vrel >= V_FRAG;
        double sample = high_speed ? col_uniform(pidx, seed, ctr) : 0.0;
        int category;
        double log_mass;
        double size_new = sample_outcome(size_i, size_j, high_speed, sample, category, log_mass);
        b.cat_count[category]++;                                                 // _record_event_work
        b.cat_log[category] += log_mass;
        numr_i = mass_before / (size_new*size_new*size_new);
        size_i = size_new;
        b.time = event_time;
        b.events++;
        double mass_after = numr_i*size_i*size_i*size_i;
        double mass_scale = std::fmax(std::fabs(mass_before), std::fabs(mass_after));
        if (!std::isfinite(size_i) || !std::isfinite(numr_i) || !(size_i > 0.0) || !(numr_i > 0.0)
            || std::fabs(mass_after - mass_before) > 2.0e-12*mass_scale)
        { b.error = 6; break; }
    }
    b.complete = true;
    size[i] = size_i;
    numr[i] = numr_i;
    rng[i] = ctr;
}

inline int thread_count(int threads)
{
#ifdef _OPENMP
    return threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
    return 1;
#endif
}


// ============================================================================================
// left-balanced k-d tree (inc/swarm/kdtree/builder_thrust.h:418-526, helpers.h:95-300,
// builder_common.h:205-240, box.h:111-125) and its closest-corner-tracking top-K query
// (traverse-cct.h:26-100, index_heap.cuh:11-177, _col_cache.cuh:15-71)
// ============================================================================================
struct kd_box { float lo[3], hi[3]; };

// widest extent; a later dimension wins ties (box.h:113-125)
inline int kd_widest(const kd_box &b)
{
    int best = 0;
    float w_best = 0.0f;
    for (int d = 0; d < 3; d++)
    {
        float w = b.hi[d] - b.lo[d];
        if (w < w_best) continue;
        w_best = w;
        best = d;
    }
    return best;
}

inline int kd_level_of(int node) { return 63 - __builtin_clzll(static_cast<unsigned long long>(node) + 1ull); }
inline int kd_levels_for(int n) { return kd_level_of(n - 1) + 1; }
inline int kd_full_nodes(int levels) { return (1 << levels) - 1; }

// nodes in the subtree rooted at `root` of a left-balanced tree of n nodes (helpers.h:185-200)
inline int kd_subtree_size(int n, int root)
{
    int levels_tree = kd_levels_for(n);
    int levels_sub = levels_tree - kd_level_of(root);
    int first = (root + 1) << (levels_sub - 1);
    int last_on_last = first + (1 << (levels_sub - 1)) - 1;
    int on_last = 1 << (levels_sub - 1);
    int missing = std::min(std::max(last_on_last - n, 0), on_last);
    return kd_full_nodes(levels_sub) - missing;
}

// position of a level-L subtree's pivot in the step-L array layout (helpers.h:214-280)
inline int kd_pivot_pos(int L, int n, int subtree)
{
    int settled = kd_full_nodes(L);
    int levels_total = kd_levels_for(n);
    int remaining = levels_total - L;
    int earlier = subtree - settled;
    int left_if_full = earlier*kd_full_nodes(remaining);
    int left_on_last_if_full = earlier*(1 << (remaining - 1));
    int total_on_last = n - kd_full_nodes(levels_total - 1);
    int missing = left_on_last_if_full - std::min(total_on_last, left_on_last_if_full);
    int seg_begin = settled + left_if_full - missing;
    int left_root = 2*subtree + 1;
    int left_size = left_root >= n ? 0 : kd_subtree_size(n, left_root);
    return seg_begin + left_size;
}

struct kd_rec { float p[3]; int idx; int dim; };

// the subtree's domain: world bounds clipped by every ancestor's split plane (builder_common.h:205-240)
inline kd_box kd_domain(int subtree, const kd_box &world, const std::vector<kd_rec> &nodes)
{
    kd_box b = world;
    int cur = subtree;
    while (cur > 0)
    {
        int parent = (cur + 1) / 2 - 1;
        int d = nodes[parent].dim;
        float split = nodes[parent].p[d];
        if (cur & 1) b.hi[d] = std::fmin(split, b.hi[d]);
        else b.lo[d] = std::fmax(split, b.lo[d]);
        cur = parent;
    }
    return b;
}


// ============================================================================================
// COL_DIAGNOSTICS: controller archive (_col_bound.cuh:219-299) and per-operator log line
// (_col_chain.cuh:26-31, 262-538)
// ============================================================================================
struct diag_bath
{
    int group, op, bath, merged;
    double duration, before, after;
    bath_result r;
};
struct diag_group
{
    unsigned long long updates = 0, coarsened = 0, refined = 0, constrained = 0;
    int overshoots = 0, persistent = 0, initial_level = 0, binding = 0, owners = 0, level = 0;
    double max_age = 0, max_growth = 0, max_activity = 0, max_ratio = 0;
    double min_request = 0, max_request = 0, initial_request = 0, final_request = 0, dt = 0;
};
struct diag_t
{
    // summary since the last reset
    int operator_count = 0, wave_count = 0, continuation_launches = 0;
    std::vector<diag_bath> baths;
    // last operator call
    unsigned long long waves = 0, owner_updates = 0, chain_blocks = 0, chain_launches = 0, finest_ticks = 0;
    double duration = 0;
    std::vector<diag_group> groups;
    long long cat_count[7] = {0, 0, 0, 0, 0, 0, 0};
    double cat_log[7] = {0, 0, 0, 0, 0, 0, 0};
};

std::string diag_controller_json(const diag_t &d)
{
    int ao = 0, dov = 0, po = 0;
    double dmin = INFINITY, dmax = 0, lmin = 1.0, mf = 0, me = 0, mt = 0, mev = 0, mg = 0, mgu = 0, mdb = 0;
    for (const auto &b : d.baths)
    {
        ao += b.r.activity_overshoot ? 1 : 0;
        dov += b.r.distribution_overshoot ? 1 : 0;
        po += b.r.persistent_overshoot ? 1 : 0;
        dmin = std::min(dmin, b.duration);
        dmax = std::max(dmax, b.duration);
        lmin = std::min(lmin, b.after);
        mf = std::max(mf, b.r.max_f);
        me = std::max(me, b.r.max_e);
        mt = std::max(mt, b.r.max_touched);
        mev = std::max(mev, b.r.max_events);
        mg = std::max(mg, b.r.max_g);
        mgu = std::max(mgu, b.r.max_g_upper);
        mdb = std::max(mdb, b.r.d_bath);
    }
    char buf[512];
    std::string o = "{\n  \"schema\": 2,\n";
    auto num = [&](const char *k, double v, bool last = false) {
        std::snprintf(buf, sizeof buf, "  \"%s\": %.17g%s\n", k, v, last ? "" : ",");
        o += buf;
    };
    num("operator_count", d.operator_count);
    num("bath_count", double(d.baths.size()));
    num("wave_count", d.wave_count);
    num("continuation_launches", d.continuation_launches);
    num("activity_overshoots", ao);
    num("distribution_overshoots", dov);
    num("persistent_overshoots", po);
    num("minimum_duration", std::isfinite(dmin) ? dmin : 0.0);
    num("maximum_duration", dmax);
    num("minimum_limit_scale", lmin);
    num("maximum_f", mf);
    num("maximum_e", me);
    num("maximum_touched", mt);
    num("maximum_events", mev);
    num("maximum_g", mg);
    num("maximum_g_upper", mgu);
    num("maximum_d_bath", mdb);
    o += "  \"baths\": [\n";
    for (size_t i = 0; i < d.baths.size(); i++)
    {
        const diag_bath &b = d.baths[i];
        std::snprintf(buf, sizeof buf,
                      "    {\"operator\": %d, \"group\": %d, \"bath\": %d, \"merged_bins\": %d, \"continuation_launches\": 0, "
                      "\"duration\": %.17g, \"limit_before\": %.17g, \"limit_after\": %.17g, \"max_f\": %.17g, "
                      "\"max_e\": %.17g, \"max_touched\": %.17g, \"max_events\": %.17g, \"max_g\": %.17g, ",
                      b.op, b.group, b.bath, b.merged, b.duration, b.before, b.after, b.r.max_f, b.r.max_e,
                      b.r.max_touched, b.r.max_events, b.r.max_g);
        o += buf;
        std::snprintf(buf, sizeof buf,
                      "\"max_g_upper\": %.17g, \"d_bath\": %.17g, \"activity_overshoot\": %s, "
                      "\"distribution_overshoot\": %s, \"persistent_overshoot\": %s}%s\n",
                      b.r.max_g_upper, b.r.d_bath, b.r.activity_overshoot ? "true" : "false",
                      b.r.distribution_overshoot ? "true" : "false", b.r.persistent_overshoot ? "true" : "false",
                      i + 1 < d.baths.size() ? "," : "");
        o += buf;
    }
    o += "  ]\n}\n";
    return o;
}

// the last operator's record without the wall-clock fields (chain, audit and scheduler seconds)
std::string diag_operator_json(const diag_t &d)
{
    char buf[1024];
    std::string o;
    std::snprintf(buf, sizeof buf,
                  "{\"duration\":%.17g,\"waves\":%llu,\"owner_updates\":%llu,\"chain_blocks\":%llu,"
                  "\"chain_launches\":%llu,\"finest_ticks\":%llu,\"groups\":[",
                  d.duration, d.waves, d.owner_updates, d.chain_blocks, d.chain_launches, d.finest_ticks);
    o += buf;
    for (size_t c = 0; c < d.groups.size(); c++)
    {
        const diag_group &g = d.groups[c];
        std::snprintf(buf, sizeof buf,
                      "%s{\"id\":%zu,\"owners\":%d,\"level\":%d,\"dt\":%.17g,\"initial_level\":%d,\"coarsened_updates\":%llu,"
                      "\"refined_updates\":%llu,\"neighbor_constrained_updates\":%llu,\"minimum_requested_dt\":%.17g,"
                      "\"maximum_requested_dt\":%.17g,\"final_requested_dt\":%.17g,\"initial_requested_dt\":%.17g,",
                      c ? "," : "", c, g.owners, g.level, g.dt, g.initial_level, g.coarsened, g.refined, g.constrained,
                      g.min_request, g.max_request, g.final_request, g.initial_request);
        o += buf;
        std::snprintf(buf, sizeof buf,
                      "\"initial_binding_constraint\":%d,\"updates\":%llu,\"overshoots\":%d,\"persistent_overshoots\":%d,"
                      "\"max_snapshot_age_at_start\":%.17g,\"max_requested_dt_ratio\":%.17g,\"max_growth\":%.17g,"
                      "\"max_predicted_activity\":%.17g}",
                      g.binding, g.updates, g.overshoots, g.persistent, g.max_age, g.max_ratio, g.max_growth,
                      g.max_activity);
        o += buf;
    }
    o += "],\"event_counts\":[";
    for (int k = 0; k < 7; k++) { std::snprintf(buf, sizeof buf, "%s%lld", k ? "," : "", d.cat_count[k]); o += buf; }
    o += "],\"event_log_mass_sums\":[";
    for (int k = 0; k < 7; k++) { std::snprintf(buf, sizeof buf, "%s%.17g", k ? "," : "", d.cat_log[k]); o += buf; }
    o += "]}";
    return o;
}

}  // namespace

// ================================================================================================
// C API
// ================================================================================================
extern "C" {

void orc_get_model_info(orc_model_info *o)
{
    o->n_k = N_K; o->dim = DIM; o->coag_kernel = COAG_KERNEL; o->n_x = N_X; o->n_y = N_Y; o->n_z = N_Z;
    o->local_groups = LOCAL_GROUPS; o->col_bin_s = COL_BIN_S; o->raw_bins = RAW_BINS;
    o->morton_leaf_target = MORTON_LEAF_TARGET; o->morton_max_level = MORTON_MAX_LEVEL; o->x_wedge = X_WEDGE;
    o->y_min = Y_MIN; o->y_max = Y_MAX; o->z_min = Z_MIN; o->z_max = Z_MAX; o->x_min = X_MIN; o->x_max = X_MAX;
    o->h_search = H_SEARCH; o->v_frag = V_FRAG; o->col_bath_max = COL_BATH_MAX; o->col_bath_eps = COL_BATH_EPS;
    o->col_bath_alpha = COL_BATH_ALPHA; o->init_smin = INIT_SMIN; o->rho_0 = RHO_0;
}

void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox10(ctr, key, out); }
double orc_uniform(uint32_t pidx, uint64_t seed, uint64_t counter) { return uniform(pidx, seed, counter); }

double orc_vrel_t(double R, double si, double sj, double h, double sig) { return vrel_t(R, si, sj, h, sig); }
double orc_vrel_pair(double y, double z, double si, double sj) { return vrel_pair(y, z, si, sj); }
double orc_ball_measure(double y, double z, double radius) { return ball_measure(y, z, radius); }
double orc_outcome_moments(double si, double sj, int high, double *mean, double *second, double *maximum)
{ return outcome_moments(si, sj, high != 0, *mean, *second, *maximum); }
double orc_sample_outcome(double si, double sj, int high, double u, int *category, double *log_mass)
{ return sample_outcome(si, sj, high != 0, u, *category, *log_mass); }
double orc_chain_rate(double y_i, double z_i, double y_j, double z_j, double size_i, double size_j,
                      double numr_j, double lambda_0, double *vrel)
{ return chain_rate(y_i, z_i, y_j, z_j, size_i, size_j, numr_j, lambda_0, *vrel); }
double orc_stokes(double y, double z, double size)
{
    double R = cyl_R(y, z), Z = cyl_Z(y, z);
    return stokes(R, Z, hg(R), size);
}
// the chain's partner pick over K rates; last_positive < 0 computes it (-1 if no rate is positive)
int orc_partner_walk(const double *rate, int k_count, double target, int last_positive)
{
    if (last_positive == -3)
    {
        last_positive = -1;
        for (int k = 0; k < k_count; k++)
            if (rate[k] > 0.0) last_positive = k;
    }
    return partner_walk(rate, k_count, target, last_positive);
}

// col_site_init, Morton variant (src/swarm/col_site_init.cu:79-118)
int orc_site_init(int n, const double *x, const double *y, const double *z, const double *size,
                  const double *numr, float *pts, float *posx, float *search_dist, uint8_t *active)
{
    int bad = 0;
    for (int i = 0; i < n; i++)
    {
        bool finite = std::isfinite(x[i]) && std::isfinite(y[i]) && std::isfinite(z[i])
            && std::isfinite(size[i]) && std::isfinite(numr[i]);
        if (!finite)
        {
            active[i] = 0;
            pts[3*i] = pts[3*i + 1] = pts[3*i + 2] = 0.0f;
            posx[i] = 0.0f;
            search_dist[i] = 0.0f;
            if (!bad) bad = i + 1;
            continue;
        }
        bool act = is_active(y[i], z[i]);
        active[i] = uint8_t(act);
        double R = cyl_R(y[i], z[i]);
        double Z = cyl_Z(y[i], z[i]);
        if (N_X == 1 && N_Z == 1)
        {
            pts[3*i] = float(R); pts[3*i + 1] = 0.0f; pts[3*i + 2] = 0.0f;
        }
        else
        {
            pts[3*i] = float(R*std::cos(x[i])); pts[3*i + 1] = float(R*std::sin(x[i])); pts[3*i + 2] = float(Z);
        }
        posx[i] = float(x[i]);
        search_dist[i] = act ? float(H_SEARCH*hg(R)*R) : 0.0f;
    }
    return bad;
}

// morton_keygen (inc/swarm/morton/morton_index.cuh:41-60)
void orc_morton_keys(int n, const float *pts, float ox, float oy, float oz, float width, int max_level,
                     int dim, uint64_t *keys)
{
    int cells = 1 << max_level;
    float scale = float(cells) / width;
    for (int i = 0; i < n; i++)
    {
        int ix = std::min(cells - 1, std::max(0, int((pts[3*i] - ox)*scale)));
        int iy = std::min(cells - 1, std::max(0, int((pts[3*i + 1] - oy)*scale)));
        int iz = (dim == 2) ? 0 : std::min(cells - 1, std::max(0, int((pts[3*i + 2] - oz)*scale)));
        keys[i] = expand3(uint32_t(ix)) | (expand3(uint32_t(iy)) << 1) | (expand3(uint32_t(iz)) << 2);
    }
}

// thrust::stable_sort_by_key contract (morton_index.cuh:187-188)
void orc_sort_perm(int n, const uint64_t *keys, int *perm)
{
    for (int i = 0; i < n; i++) perm[i] = i;
    std::stable_sort(perm, perm + n, [&](int a, int b) { return keys[a] < keys[b]; });
}

// morton_index::build level loop (morton_index.cuh:190-225; morton_counts 94-116, morton_expand 119-145)
int orc_hierarchy_build(int n, const uint64_t *keys, float ox, float oy, float oz, float width,
                        int dim, int leaf_target, int max_level, int cap, float *lower, float *nwidth,
                        int *begin, int *count, int *child, int *child_count)
{
    if (cap < 1) return -1;
    auto cell = [&](int id, int b, int c, float lx, float ly, float lz, float w) {
        lower[3*id] = lx; lower[3*id + 1] = ly; lower[3*id + 2] = lz;
        nwidth[id] = w; begin[id] = b; count[id] = c; child_count[id] = 0;
        for (int code = 0; code < 8; code++) child[8*id + code] = -1;
    };
    cell(0, 0, n, ox, oy, oz, width);
    int node_count = 1, first = 0, frontier = 1;
    for (int level = 0; level <= max_level && frontier; ++level)
    {
        std::vector<uint64_t> offsets(size_t(frontier) + 1, 0);
        for (int idx = 0; idx < frontier; idx++)
        {
            int id = first + idx, cc = 0;
            if (count[id] > leaf_target && level < max_level)
            {
                int b = begin[id], e = b + count[id], shift = 3*(max_level - level - 1);
                for (int code = 1; code <= (1 << dim); ++code)
                {
                    int next = child_end(keys, b, e, shift, code);
                    cc += next > b;
                    b = next;
                }
            }
            child_count[id] = cc;
            offsets[idx + 1] = offsets[idx] + uint64_t(cc);
        }
        uint64_t next_count = offsets[frontier];
        if (!next_count) break;
        if (uint64_t(node_count) + next_count > uint64_t(cap)) return -1;
        for (int idx = 0; idx < frontier; idx++)
        {
            int id = first + idx;
            if (!child_count[id]) continue;
            int b = begin[id], e = b + count[id];
            int slot = node_count + int(offsets[idx]);
            int shift = 3*(max_level - level - 1);
            float w = 0.5f*nwidth[id];
            for (int code = 0; code < (1 << dim); ++code)
            {
                int next = child_end(keys, b, e, shift, code + 1);
                if (next > b)
                {
                    float lx = lower[3*id], ly = lower[3*id + 1], lz = lower[3*id + 2];
                    if (code & 1) lx += w;
                    if (code & 2) ly += w;
                    if (dim == 3 && (code & 4)) lz += w;
                    cell(slot, b, next - b, lx, ly, lz, w);
                    child[8*id + code] = slot++;
                }
                b = next;
            }
        }
        first = node_count;
        node_count += int(next_count);
        frontier = int(next_count);
    }
    return node_count;
}

// brute-force exact top-K (val/swarm/mod/test_knn/knn_benchmark.cu:508-545 contract)
void orc_topk_brute(int n, const float *pts, const uint8_t *active, int nq, const int *query_idx,
                    const float *query_pts, const float *query_q, int k, int *out_codes, float *out_d2, int threads)
{
    (void)query_idx;
    #pragma omp parallel for schedule(dynamic, 16) num_threads(thread_count(threads))
    for (int q = 0; q < nq; q++)
    {
        std::vector<std::pair<float, int>> cand;
        float qd = query_q[q];
        float qd2 = qd*qd;
        for (int j = 0; j < n; j++)
        {
            if (!active[j]) continue;
            float d2 = point_d2(query_pts + 3*q, pts + 3*j);
            if (d2 <= qd2) cand.emplace_back(d2, 3*j);
        }
        int take = std::min<int>(k, int(cand.size()));
        std::partial_sort(cand.begin(), cand.begin() + take, cand.end(),
            [](const std::pair<float, int> &a, const std::pair<float, int> &b) { return nb_less(a.first, a.second, b.first, b.second); });
        for (int s = 0; s < k; s++)
        {
            out_codes[size_t(q)*k + s] = s < take ? cand[s].second : -1;
            out_d2[size_t(q)*k + s] = s < take ? cand[s].first : std::numeric_limits<float>::infinity();
        }
    }
}

// ---- azimuthal wedge (periodic images) ----
// _get_image_dist_min (inc/swarm/param_grid.cuh:86-89): the closest two images of one particle get
float orc_image_dist_min()
{
    const float x_min = float(X_MIN), x_max = float(X_MAX);
    auto sinf1 = [](float a) { return float(std::sin(double(a))); };
    return 2.0f*float(Y_MIN)*std::fmin(sinf1(float(Z_MIN)), sinf1(float(Z_MAX)))*std::fabs(sinf1(0.5f*(x_max - x_min)));
}

// _get_morton_seam_dist (inc/swarm/morton/morton_ghost.cuh:28-35)
static float seam_dist(const float *p, float off)
{
    float R = std::hypot(p[0], p[1]);
    return (std::cos(off) >= 0.0f) ? R*std::fabs(std::sin(off)) : R;
}

// morton_gcount + exclusive scan + morton_gwrite (morton_ghost.cuh:38-86, 122-153): physical records
// 0..n-1 with codes 3i, then per particle in index order its image 2 (rotated by +width, near x_min)
// and image 1 (rotated by -width, near x_max). rec and code hold room for 3n; returns the count.
int orc_wedge_records(int n, const float *pts, const float *posx, float q_max, float *rec, int *code)
{
    const float x_min = float(X_MIN), x_max = float(X_MAX);
    const float width = x_max - x_min;
    // sin and cos of the float width, each rounded once to float
    const float sw = float(std::sin(double(width))), cw = float(std::cos(double(width)));
    for (int i = 0; i < n; i++)
    {
        rec[3*i] = pts[3*i]; rec[3*i + 1] = pts[3*i + 1]; rec[3*i + 2] = pts[3*i + 2];
        code[i] = 3*i;
    }
    int g = n;
    for (int i = 0; i < n; i++)
    {
        const float *p = pts + 3*size_t(i);
        if (seam_dist(p, posx[i] - x_min) <= q_max)
        {
            rec[3*g] = cw*p[0] - sw*p[1]; rec[3*g + 1] = sw*p[0] + cw*p[1]; rec[3*g + 2] = p[2];
            code[g++] = 3*i + 2;
        }
        if (seam_dist(p, x_max - posx[i]) <= q_max)
        {
            rec[3*g] = cw*p[0] + sw*p[1]; rec[3*g + 1] = -sw*p[0] + cw*p[1]; rec[3*g + 2] = p[2];
            code[g++] = 3*i + 1;
        }
    }
    return g;
}

// exact top-K over coded records by brute force; with dedup, _morton_ghost_topk (morton_query.cuh:62-140):
// the 3K nearest records, each physical particle kept at its nearest image, the first K of those
void orc_topk_records(int nr, const float *rec, const int *code, const uint8_t *active, int nq, const float *query_pts,
                      const float *query_q, int k, int dedup, int *out_codes, float *out_d2, int threads)
{
    #pragma omp parallel for schedule(dynamic, 16) num_threads(thread_count(threads))
    for (int q = 0; q < nq; q++)
    {
        std::vector<std::pair<float, int>> cand;
        float qd = query_q[q];
        float qd2 = qd*qd;
        for (int r = 0; r < nr; r++)
        {
            if (!active[code[r] / 3]) continue;
            float d2 = point_d2(query_pts + 3*size_t(q), rec + 3*size_t(r));
            if (d2 <= qd2) cand.emplace_back(d2, code[r]);
        }
        auto less = [](const std::pair<float, int> &a, const std::pair<float, int> &b) {
            return nb_less(a.first, a.second, b.first, b.second);
        };
        std::sort(cand.begin(), cand.end(), less);
        if (dedup)
        {
            if (cand.size() > size_t(3*k)) cand.resize(size_t(3*k));
            std::vector<std::pair<float, int>> kept;
            for (size_t a = 0; a < cand.size(); a++)
            {
                bool dup = false;
                for (size_t b = 0; b < a; b++) dup = dup || cand[b].second / 3 == cand[a].second / 3;
                if (!dup) kept.push_back(cand[a]);
            }
            cand.swap(kept);
        }
        int take = std::min<int>(k, int(cand.size()));
        for (int s = 0; s < k; s++)
        {
            out_codes[size_t(q)*k + s] = s < take ? cand[s].second : -1;
            out_d2[size_t(q)*k + s] = s < take ? cand[s].first : std::numeric_limits<float>::infinity();
        }
    }
}

// exact top-K through the Morton hierarchy (traversal contract of _morton_topk, morton_index.cuh:493-679)
void orc_topk_tree(int n, const float *pts, const uint8_t *active, const int *sorted_idx, int node_count,
                   const float *lower, const float *nwidth, const int *begin, const int *count, const int *child,
                   const int *child_count, int dim, int max_level, int nq, const int *query_idx, int k,
                   const float *search_dist, int *out_codes, float *out_d2, int threads)
{
    (void)n; (void)node_count;
    #pragma omp parallel for schedule(dynamic, 64) num_threads(thread_count(threads))
    for (int q = 0; q < nq; q++)
    {
        int i = query_idx[q];
        const float *qp = pts + 3*size_t(i);
        float qd = search_dist[i];
        float qd2 = qd*qd;
        std::vector<float> d(k, std::numeric_limits<float>::infinity());
        std::vector<int> c(k, INT_MAX);
        std::vector<int> stack;
        stack.push_back(0);
        while (!stack.empty())
        {
            int id = stack.back(); stack.pop_back();
            float cutoff = std::fmin(qd2, d[k - 1]);
            if (node_d2(qp, lower + 3*id, nwidth[id], dim, max_level) > cutoff) continue;
            if (child_count[id] == 0)
            {
                for (int s = begin[id]; s < begin[id] + count[id]; s++)
                {
                    int j = sorted_idx[s];
                    if (!active[j]) continue;
                    float d2 = point_d2(qp, pts + 3*size_t(j));
                    int code = 3*j;
                    if (!(d2 <= qd2) || !nb_less(d2, code, d[k - 1], c[k - 1])) continue;
                    int pos = k - 1;
                    while (pos > 0 && nb_less(d2, code, d[pos - 1], c[pos - 1])) { d[pos] = d[pos - 1]; c[pos] = c[pos - 1]; pos--; }
                    d[pos] = d2; c[pos] = code;
                }
                continue;
            }
            int ids[8]; float dd[8]; int nv = 0;
            for (int code = 0; code < (dim == 2 ? 4 : 8); code++)
            {
                int ch = child[8*id + code];
                if (ch < 0) continue;
                float cd = node_d2(qp, lower + 3*ch, nwidth[ch], dim, max_level);
                if (cd > cutoff) continue;
                int p = nv;
                while (p > 0 && dd[p - 1] > cd) { dd[p] = dd[p - 1]; ids[p] = ids[p - 1]; p--; }
                dd[p] = cd; ids[p] = ch; nv++;
            }
            for (int p = nv - 1; p >= 0; p--) stack.push_back(ids[p]);
        }
        for (int s = 0; s < k; s++)
        {
            out_codes[size_t(q)*k + s] = c[s] == INT_MAX ? -1 : c[s];
            out_d2[size_t(q)*k + s] = d[s];
        }
    }
}

// k-d tree build over all n sites (inactive ones included, as the reference's N_T nodes):
// out_idx[node] = particle index, out_dim[node] = split dimension, world6 = {lo xyz, hi xyz}.
// Each level stable-sorts (tag, coordinate along the point's dimension); the reference's sort is
// not stable, so its layout for equal coordinates is unspecified and this one is the stable choice.
void orc_kd_build(int n, const float *pts, int *out_idx, int *out_dim, float *world6)
{
    if (n < 1) return;
    std::vector<kd_rec> nodes(n);
    kd_box world;
    for (int d = 0; d < 3; d++) { world.lo[d] = pts[d]; world.hi[d] = pts[d]; }
    for (int i = 0; i < n; i++)
    {
        for (int d = 0; d < 3; d++)
        {
            nodes[i].p[d] = pts[3*size_t(i) + d];
            world.lo[d] = std::fmin(world.lo[d], nodes[i].p[d]);
            world.hi[d] = std::fmax(world.hi[d], nodes[i].p[d]);
        }
        nodes[i].idx = i;
    }
    int dim0 = kd_widest(world);
    for (auto &r : nodes) r.dim = dim0;
    std::vector<uint32_t> tag(n, 0u);
    std::vector<int> order(n);
    auto sort_level = [&]() {
        for (int i = 0; i < n; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            if (tag[a] != tag[b]) return tag[a] < tag[b];
            int d = nodes[a].dim;
            return nodes[a].p[d] < nodes[b].p[d];
        });
        std::vector<kd_rec> nn(n);
        std::vector<uint32_t> nt(n);
        for (int i = 0; i < n; i++) { nn[i] = nodes[order[i]]; nt[i] = tag[order[i]]; }
        nodes.swap(nn);
        tag.swap(nt);
    };
    const int deepest = kd_levels_for(n) - 1;
    for (int L = 0; L < deepest; L++)
    {
        sort_level();
        const int settled = kd_full_nodes(L);
        std::vector<int> new_dim(n);
        for (int g = 0; g < n; g++) new_dim[g] = nodes[g].dim;
        for (int g = settled; g < n; g++)
        {
            int subtree = int(tag[g]);
            kd_box b = kd_domain(subtree, world, nodes);
            int pivot = kd_pivot_pos(L, n, subtree);
            int pd = nodes[pivot].dim;
            float pc = nodes[pivot].p[pd];
            if (g < pivot) { subtree = 2*subtree + 1; b.hi[pd] = pc; }
            else if (g > pivot) { subtree = 2*subtree + 2; b.lo[pd] = pc; }
            if (g != pivot) new_dim[g] = kd_widest(b);
            tag[g] = uint32_t(subtree);
        }
        for (int g = 0; g < n; g++) nodes[g].dim = new_dim[g];
    }
    sort_level();
    for (int i = 0; i < n; i++) { out_idx[i] = nodes[i].idx; out_dim[i] = nodes[i].dim; }
    for (int d = 0; d < 3; d++) { world6[d] = world.lo[d]; world6[3 + d] = world.hi[d]; }
}

// top-K of the queries over the tree: a max-heap of (d^2 bits << 32 | code) keys initialised to
// (q^2, 0xffffffff), culled at nextafter(root d^2), children visited near side first with the
// far side's closest corner tracked on a stack. Output rows are in heap order; codes 3*j, -1 empty.
// With n_phys > 0 (an azimuthal wedge) the tree holds n_phys physical points and their images 1 and 2
// at [n_phys, 2 n_phys) and [2 n_phys, 3 n_phys): node record r is particle r % n_phys at image
// r / n_phys, code 3*(r % n_phys) + image, active by its particle. A query whose two images can share
// its ball (image_dist_min <= 2 q, col_cache_get in inc/swarm/_col_cache.cuh:38-41) keeps one entry
// per particle: a candidate of a particle already in the heap replaces that entry only when smaller
// (idx_old_heap::processCandidate, inc/swarm/kdtree/index_heap.cuh:112-175).
static void kd_query(int n, const float *pts, const uint8_t *active, const int *node_idx, const int *node_dim,
                     const float *world6, int nq, const int *query_idx, int k, const float *search_dist,
                     int *out_codes, float *out_d2, int threads, int n_phys, float image_dist_min)
{
    kd_box world;
    for (int d = 0; d < 3; d++) { world.lo[d] = world6[d]; world.hi[d] = world6[3 + d]; }
    #pragma omp parallel for schedule(dynamic, 64) num_threads(thread_count(threads))
    for (int qi = 0; qi < nq; qi++)
    {
        const int i = query_idx[qi];
        const float *qp = pts + 3*size_t(i);
        const float qd = search_dist[i];
        auto enc = [](float d2, uint32_t code) {
            uint32_t b;
            std::memcpy(&b, &d2, 4);
            return (uint64_t(b) << 32) | code;
        };
        auto dec = [](uint64_t key) {
            uint32_t b = uint32_t(key >> 32);
            float f;
            std::memcpy(&f, &b, 4);
            return f;
        };
        std::vector<uint64_t> heap(k, enc(qd*qd, 0xffffffffu));
        const bool dedup = n_phys > 0 && !(image_dist_min < 0.0f || image_dist_min > 2.0f*qd);
        auto cull = [&]() { return std::nextafter(dec(heap[0]), std::numeric_limits<float>::infinity()); };
        auto process = [&](int node, float d2) {
            int r = node_idx[node];
            int j = n_phys > 0 ? r % n_phys : r;
            int image = n_phys > 0 ? r / n_phys : 0;
            if (!active[j]) return;
            uint64_t cand = enc(d2, uint32_t(3*j + image));
            int slot = -1;
            if (dedup)
                for (int s = 0; s < k; s++)
                {
                    uint32_t c = uint32_t(heap[s]);
                    if (c != 0xffffffffu && int(c / 3) == j) { slot = s; break; }
                }
            if (slot >= 0)
            {
                if (cand >= heap[slot]) return;
            }
            else
            {
                if (cand >= heap[0]) return;
                slot = 0;
            }
            while (true)
            {
                int child = 2*slot + 1;
                if (child >= k) { heap[slot] = cand; break; }
                uint64_t cv = heap[child];
                if (child + 1 < k && heap[child + 1] > cv) { cv = heap[child + 1]; ++child; }
                if (cv < cand) { heap[slot] = cand; break; }
                heap[slot] = cv;
                slot = child;
            }
        };
        float cull_d2 = cull();
        float corner[3];
        for (int d = 0; d < 3; d++) corner[d] = std::fmin(std::fmax(qp[d], world.lo[d]), world.hi[d]);
        struct entry { int node; float c[3]; };
        std::vector<entry> stack;
        bool run = point_d2(corner, qp) <= cull_d2;
        int node = 0;
        while (run)
        {
            if (node >= n)
            {
                bool found = false;
                while (!stack.empty())
                {
                    entry e = stack.back();
                    stack.pop_back();
                    for (int d = 0; d < 3; d++) corner[d] = e.c[d];
                    if (point_d2(corner, qp) >= cull_d2) continue;
                    node = e.node;
                    found = true;
                    break;
                }
                if (!found) break;
            }
            const float *np = pts + 3*size_t(node_idx[node]);
            process(node, point_d2(np, qp));
            cull_d2 = cull();
            int dim = node_dim[node];
            bool left_close = qp[dim] < np[dim];
            int lc = 2*node + 1, rc = lc + 1;
            int far = left_close ? rc : lc;
            float fc[3] = {corner[0], corner[1], corner[2]};
            fc[dim] = np[dim];
            if (far < n && point_d2(fc, qp) < cull_d2)
            {
                entry e{far, {fc[0], fc[1], fc[2]}};
                stack.push_back(e);
            }
            node = left_close ? lc : rc;
        }
        for (int s2 = 0; s2 < k; s2++)
        {
            uint32_t code = uint32_t(heap[s2]);
            out_codes[size_t(qi)*k + s2] = code == 0xffffffffu ? -1 : int(code);
            out_d2[size_t(qi)*k + s2] = dec(heap[s2]);
        }
    }
}

void orc_kd_query(int n, const float *pts, const uint8_t *active, const int *node_idx, const int *node_dim,
                  const float *world6, int nq, const int *query_idx, int k, const float *search_dist,
                  int *out_codes, float *out_d2, int threads)
{
    kd_query(n, pts, active, node_idx, node_dim, world6, nq, query_idx, k, search_dist, out_codes, out_d2, threads,
             0, -1.0f);
}

// the k-d search of an azimuthal wedge: pts holds 3 n_phys points (orc_kd_images), query_idx and
// search_dist index physical particles
void orc_kd_query_wedge(int n_phys, const float *pts, const uint8_t *active, const int *node_idx,
                        const int *node_dim, const float *world6, int nq, const int *query_idx, int k,
                        const float *search_dist, int *out_codes, float *out_d2, int threads)
{
    kd_query(3*n_phys, pts, active, node_idx, node_dim, world6, nq, query_idx, k, search_dist, out_codes, out_d2,
             threads, n_phys, orc_image_dist_min());
}

// col_site_init's wedge images for the k-d tree (src/swarm/col_site_init.cu:42-62): out holds 3n points,
// the physical ones, then every point rotated by -width (image 1), then by +width (image 2); the width
// is float(X_MAX - X_MIN), its sine and cosine each rounded once from double
void orc_kd_images(int n, const float *pts, float *out)
{
    const float width = float(X_MAX - X_MIN);
    const float sw = float(std::sin(double(width))), cw = float(std::cos(double(width)));
    for (int i = 0; i < n; i++)
    {
        const float *p = pts + 3*size_t(i);
        for (int d = 0; d < 3; d++) out[3*size_t(i) + d] = p[d];
        for (int image = 0; image < 2; image++)
        {
            const float sa = image == 0 ? -sw : sw;
            float *o = out + 3*(size_t(i) + size_t(image + 1)*n);
            o[0] = cw*p[0] - sa*p[1];
            o[1] = sa*p[0] + cw*p[1];
            o[2] = p[2];
        }
    }
}

void orc_space_bin(int n, const double *x, const double *y, const double *z, int *spatial)
{
    for (int i = 0; i < n; i++) spatial[i] = space_bin(x[i], y[i], z[i]);
}

int orc_change_limit(double A, double B, double eps, double horizon, double *duration, int *reason)
{
    try { auto r = change_limit(A, B, eps, horizon); *duration = r.duration; *reason = r.reason; return 0; }
    catch (const std::exception &) { return 1; }
}

int orc_build_binmap(const int *raw_count, int raw_bins, int *raw_to_merged)
{
    std::vector<int> rc(raw_count, raw_count + raw_bins), map;
    int merged = build_binmap(rc, map);
    for (int i = 0; i < raw_bins; i++) raw_to_merged[i] = map[i];
    return merged;
}

int orc_choose_bath(const double *bins6, int n, double remaining, double limit_scale, double *duration, int *binding)
{
    std::vector<rate_bin> bins(n);
    for (int b = 0; b < n; b++)
    {
        bins[b].mass = bins6[6*b]; bins[b].weighted_rate = bins6[6*b + 1]; bins[b].weighted_change = bins6[6*b + 2];
        bins[b].weighted_second = bins6[6*b + 3]; bins[b].owner_count = int(bins6[6*b + 4]); bins[b].invalid_count = int(bins6[6*b + 5]);
    }
    try { *duration = choose_bath(bins, remaining, limit_scale, binding); return 0; }
    catch (const std::exception &) { return 1; }
}

int orc_finish_bath(const double *a, int n, double *state4, double *result10)
{
    std::vector<audit_bin> bins(n);
    for (int b = 0; b < n; b++)
    {
        const double *v = a + 15*b;
        auto &o = bins[b];
        o.mass = v[0]; o.predicted_f = v[1]; o.predicted_e = v[2]; o.predicted_var_f = v[3]; o.predicted_var_e = v[4];
        o.predicted_g = v[5]; o.predicted_var_g = v[6]; o.maximum_weight = v[7]; o.maximum_g_jump = v[8];
        o.touched = v[9]; o.event_weight = v[10]; o.growth = v[11]; o.end_mass = v[12];
        o.owner_count = int(v[13]); o.invalid_count = int(v[14]);
    }
    bath_state s;
    s.limit_scale = state4[0]; s.activity_streak = int(state4[1]); s.quiet_streak = int(state4[2]); s.floor_streak = int(state4[3]);
    try
    {
        bath_result r = finish_bath(bins, s);
        state4[0] = s.limit_scale; state4[1] = s.activity_streak; state4[2] = s.quiet_streak; state4[3] = s.floor_streak;
        double out[10] = {r.max_f, r.max_e, r.max_touched, r.max_events, r.max_g, r.max_g_upper, r.d_bath,
                          double(r.activity_overshoot), double(r.distribution_overshoot), double(r.persistent_overshoot)};
        for (int i = 0; i < 10; i++) result10[i] = out[i];
        return 0;
    }
    catch (const std::exception &) { return 1; }
}

int orc_schedule_run(int n, double horizon, const double *requested, const uint32_t *edges, int max_waves,
                     uint64_t *ticks, int *due_counts, int *levels_out)
{
    try
    {
        std::vector<double> req(requested, requested + n);
        std::vector<uint32_t> e(edges, edges + size_t(n)*((n + 31) / 32));
        schedule s(horizon, req, e);
        for (int c = 0; c < n; c++) levels_out[c] = s.level[c];
        std::vector<bool> passed(n, true);
        int waves = 0;
        while (s.time() < s.end)
        {
            if (waves >= max_waves) return -1;
            uint64_t tick = s.time();
            auto groups = s.due(tick);
            if (tick != 0) s.adapt(tick, groups, req, passed);
            ticks[waves] = tick;
            due_counts[waves] = int(groups.size());
            waves++;
            s.advance(groups);
        }
        return waves;
    }
    catch (const std::exception &) { return -1; }
}

// local_schedule driven with a request and an audit verdict per group and wave (test access)
int orc_schedule_adaptive(int n, double horizon, const double *requested, const uint32_t *edges, int max_waves,
                          const double *wave_requested, const uint8_t *wave_passed, uint64_t *ticks, int *due_counts,
                          int *wave_levels)
{
    try
    {
        std::vector<double> req(requested, requested + n);
        std::vector<uint32_t> e(edges, edges + size_t(n)*((n + 31) / 32));
        schedule s(horizon, req, e);
        std::vector<bool> passed(n, true);
        int waves = 0;
        while (s.time() < s.end)
        {
            if (waves >= max_waves) return -1;
            uint64_t tick = s.time();
            auto groups = s.due(tick);
            if (tick != 0)
            {
                for (int c = 0; c < n; c++)
                {
                    req[c] = wave_requested[size_t(waves)*n + c];
                    passed[c] = wave_passed[size_t(waves)*n + c] != 0;
                }
                s.adapt(tick, groups, req, passed);
            }
            ticks[waves] = tick;
            due_counts[waves] = int(groups.size());
            for (int c = 0; c < n; c++) wave_levels[size_t(waves)*n + c] = s.level[c];
            waves++;
            s.advance(groups);
        }
        return waves;
    }
    catch (const std::exception &) { return -1; }
}

void *orc_diag_new(void) { return new diag_t(); }
void orc_diag_free(void *d) { delete static_cast<diag_t *>(d); }
// the controller archive's reset at output boundaries (swarm_runtime.cu:491-496)
void orc_diag_reset(void *d)
{
    diag_t &x = *static_cast<diag_t *>(d);
    x.operator_count = x.wave_count = x.continuation_launches = 0;
    x.baths.clear();
}
// JSON text: 0 = controller archive (schema 2), 1 = last operator record; returns the length
int orc_diag_json(void *d, int which, char *buf, int len)
{
    std::string s = which == 0 ? diag_controller_json(*static_cast<diag_t *>(d))
                               : diag_operator_json(*static_cast<diag_t *>(d));
    if (buf && len > 0)
    {
        std::strncpy(buf, s.c_str(), size_t(len) - 1);
        buf[len - 1] = 0;
    }
    return int(s.size());
}

int orc_ctl_size(void) { return int(sizeof(ctl_t)); }
void orc_ctl_init(void *p)
{
    ctl_t *c = static_cast<ctl_t *>(p);
    for (int g = 0; g < LOCAL_GROUPS; g++) { c->state[g] = bath_state{}; c->lower[g] = 0.0; c->upper[g] = 0.0; }
}

void orc_bath_rates(int n, const double *y, const double *z, const double *size, const double *numr,
                    const uint8_t *active, const int *nbr, const double *measure, double lambda_0,
                    double *cached4, int threads)
{
    #pragma omp parallel num_threads(thread_count(threads))
    {
        pair_work w{std::vector<double>(N_K), std::vector<double>(N_K), std::vector<double>(N_K),
                    std::vector<double>(N_K), std::vector<double>(N_K), {}, {}};
        std::vector<double> tmp;
        #pragma omp for schedule(static)
        for (int i = 0; i < n; i++)
        {
            owner_pairs(i, size[i], y, z, nbr, measure, size, numr, lambda_0, true, active, w);
            owner_totals t = reduce_pairs(w, tmp);
            cached4[4*size_t(i)] = t.valid ? t.rate : -1.0;
            cached4[4*size_t(i) + 1] = t.j1;
            cached4[4*size_t(i) + 2] = t.j2;
            cached4[4*size_t(i) + 3] = t.jmax;
        }
    }
}

// every owner's K pair terms at the bath start (col_bath_rate.cu:30-58): n x N_K rates in slot order
void orc_pair_rates(int n, const double *y, const double *z, const double *size, const double *numr,
                    const uint8_t *active, const int *nbr, const double *measure, double lambda_0, double *rates,
                    int threads)
{
    #pragma omp parallel num_threads(thread_count(threads))
    {
        pair_work w{std::vector<double>(N_K), std::vector<double>(N_K), std::vector<double>(N_K),
                    std::vector<double>(N_K), std::vector<double>(N_K), {}, {}};
        #pragma omp for schedule(static)
        for (int i = 0; i < n; i++)
        {
            owner_pairs(i, size[i], y, z, nbr, measure, size, numr, lambda_0, true, active, w);
            for (int k = 0; k < N_K; k++) rates[size_t(i)*N_K + k] = w.rate[k];
        }
    }
}

// evolve_local_collisions (inc/swarm/_col_chain.cuh:108-539) on the host, particle order
int orc_operator(int n, const double *y, const double *z, double *size, double *numr, const uint8_t *active,
                 const int *spatial, const int *nbr, const double *measure, uint64_t *rng, uint64_t seed,
                 double duration, double total_dust_mass, void *ctl_ptr, int threads, orc_stats *stats,
                 char *err, int err_len, void *diag_ptr)
{
    ctl_t &ctl = *static_cast<ctl_t *>(ctl_ptr);
    diag_t *dg = static_cast<diag_t *>(diag_ptr);
    std::memset(stats, 0, sizeof(*stats));
    const int nth = thread_count(threads);
    try
    {
        // owner lists and the dependency graph (_col_chain.cuh:161-180, col_dep_graph.cu:43-59)
        std::vector<std::vector<int>> owners(LOCAL_GROUPS);
        for (int i = 0; i < n; ++i) owners[spatial[i]].push_back(i);
        std::vector<uint32_t> edges(size_t(LOCAL_GROUPS)*LOCAL_WORDS, 0);
        for (int i = 0; i < n; i++)
        {
            if (!active[i]) continue;
            for (int k = 0; k < N_K; k++)
            {
                int code = nbr[size_t(i)*N_K + k];
                if (code < 0) continue;
                int c = spatial[code / 3];
                edges[size_t(spatial[i])*LOCAL_WORDS + c / 32] |= 1u << (c % 32);
            }
        }
        std::vector<double> size_old(n), numr_old(n), cached(size_t(4)*n), rate(n);
        std::vector<owner_bath> bath(n);
        const double lambda0 = n / double(N_K) / total_dust_mass;
        std::vector<int> counts(RAW_BINS), binmap(RAW_BINS);
        std::vector<rate_bin> ratebin(RAW_BINS);
        std::vector<audit_bin> audit(RAW_BINS);
        std::vector<int> ids;

        auto initialize = [&](const std::vector<int> &list) {                       // col_bath_init, col_comp_zero
            for (int i : list) { size_old[i] = size[i]; numr_old[i] = numr[i]; bath[i] = owner_bath{}; }
        };
        auto rates_and_bins = [&](const std::vector<int> &list) {                   // _col_chain.cuh:205-243
            std::vector<double> smin(LOCAL_GROUPS, INFINITY), smax(LOCAL_GROUPS, 0.0);
            for (int i : list)
            {
                double s = size[i];
                if (!active[i] || !(s > 0) || !std::isfinite(s)) continue;
                smin[spatial[i]] = std::min(smin[spatial[i]], s);
                smax[spatial[i]] = std::max(smax[spatial[i]], s);
            }
            for (int g = 0; g < LOCAL_GROUPS; g++)
                if (smax[g] != 0) { ctl.lower[g] = 0.5*smin[g]; ctl.upper[g] = 8.0*smax[g]; }
            int m = int(list.size());
            #pragma omp parallel num_threads(nth)
            {
                pair_work w{std::vector<double>(N_K), std::vector<double>(N_K), std::vector<double>(N_K),
                            std::vector<double>(N_K), std::vector<double>(N_K), {}, {}};
                std::vector<double> tmp;
                #pragma omp for schedule(static)
                for (int s = 0; s < m; s++)
                {
                    int i = list[s];
                    owner_pairs(i, size_old[i], y, z, nbr, measure, size_old.data(), numr_old.data(), lambda0, true, active, w);
                    owner_totals t = reduce_pairs(w, tmp);
                    cached[4*size_t(i)] = t.valid ? t.rate : -1.0;
                    cached[4*size_t(i) + 1] = t.j1;
                    cached[4*size_t(i) + 2] = t.j2;
                    cached[4*size_t(i) + 3] = t.jmax;
                    rate[i] = t.rate;
                }
            }
            stats->pair_evals += (long long)m*N_K;
            std::fill(counts.begin(), counts.end(), 0);
            for (int i : list)
            {
                if (!active[i]) continue;
                counts[spatial[i]*COL_BIN_S + size_bin(ctl, size[i], spatial[i])]++;
            }
            int merged = build_binmap(counts, binmap);
            std::fill(ratebin.begin(), ratebin.end(), rate_bin{});
            for (int i : list)
            {
                if (!active[i]) continue;
                int b = binmap[spatial[i]*COL_BIN_S + size_bin(ctl, size[i], spatial[i])];
                double weight = numr[i]*grain_mass(size[i]);
                ratebin[b].owner_count++;
                double rc = rate[i], ch = cached[4*size_t(i) + 1], se = cached[4*size_t(i) + 2];
                if (!std::isfinite(weight) || !(weight > 0.0) || !std::isfinite(rc) || rc < 0.0
                    || !std::isfinite(ch) || ch < 0 || !std::isfinite(se) || se < 0)
                { ratebin[b].invalid_count++; continue; }
                ratebin[b].mass += weight;
                ratebin[b].weighted_rate += weight*rc;
                ratebin[b].weighted_change += weight*ch;
                ratebin[b].weighted_second += weight*se;
            }
            return merged;
        };
        auto bin_end = [&](int c, int merged) { return c + 1 < LOCAL_GROUPS ? binmap[(c + 1)*COL_BIN_S] : merged; };
        auto requested_step = [&](int c, int merged, int *binding = nullptr) {
            int first = binmap[c*COL_BIN_S], last = bin_end(c, merged);
            std::vector<rate_bin> slice(ratebin.begin() + first, ratebin.begin() + last);
            return choose_bath(slice, duration, ctl.state[c].limit_scale, binding);
        };

        std::vector<int> all(n);
        for (int i = 0; i < n; i++) all[i] = i;
        initialize(all);
        int merged = rates_and_bins(all);
        std::vector<double> requested(LOCAL_GROUPS);
        std::vector<diag_group> dgs(LOCAL_GROUPS);
        for (int c = 0; c < LOCAL_GROUPS; ++c) requested[c] = requested_step(c, merged, &dgs[c].binding);
        schedule sched(duration, requested, edges);
        std::vector<double> steps(LOCAL_GROUPS);
        for (int c = 0; c < LOCAL_GROUPS; ++c) steps[c] = sched.seconds(sched.step[c]);
        std::vector<bool> passed(LOCAL_GROUPS, true);
        std::vector<double> current = requested;
        double dt_col = duration;
        std::vector<double> published(LOCAL_GROUPS, 0.0);
        std::vector<long long> cat_count(dg ? size_t(n)*7 : 0, 0);
        std::vector<double> cat_log(dg ? size_t(n)*7 : 0, 0.0);
        int op_index = 0, bath_index = 0;
        if (dg)
        {
            op_index = ++dg->operator_count;
            dg->waves = dg->owner_updates = dg->chain_blocks = dg->chain_launches = 0;
            dg->duration = duration;
            for (int c = 0; c < LOCAL_GROUPS; ++c)
            {
                dgs[c].initial_level = sched.level[c];
                dgs[c].min_request = dgs[c].max_request = dgs[c].initial_request = requested[c];
            }
        }
        while (sched.time() < sched.end)
        {
            uint64_t tick = sched.time();
            auto groups = sched.due(tick);
            double time = sched.seconds(tick);
            ids.clear();
            for (int c : groups)
            {
                published[c] = time;
                ids.insert(ids.end(), owners[c].begin(), owners[c].end());
            }
            int count = int(ids.size());
            if (count == 0) { sched.advance(groups); continue; }
            if (tick != 0)
            {
                initialize(ids);
                merged = rates_and_bins(ids);
                for (int c : groups)
                {
                    current[c] = requested_step(c, merged);
                    dgs[c].min_request = std::min(dgs[c].min_request, current[c]);
                    dgs[c].max_request = std::max(dgs[c].max_request, current[c]);
                }
                std::vector<int> previous = sched.level;
                sched.adapt(tick, groups, current, passed);
                for (int c : groups)
                {
                    steps[c] = sched.seconds(sched.step[c]);
                    dgs[c].coarsened += sched.level[c] < previous[c];
                    dgs[c].refined += sched.level[c] > previous[c];
                    dgs[c].constrained += steps[c] > current[c];
                }
            }
            for (int c : groups)
            {
                if (owners[c].empty()) continue;
                for (int d = 0; d < LOCAL_GROUPS; ++d)
                    if (edges[size_t(c)*LOCAL_WORDS + d / 32] & (1u << (d % 32)))
                        dgs[c].max_age = std::max(dgs[c].max_age, time - published[d]);
                dgs[c].max_ratio = std::max(dgs[c].max_ratio, steps[c] / requested_step(c, merged));
            }
            // screen and chain (col_skip_scan + col_chain_run)
            int error = 0;
            long long events = 0, evals = 0;
            #pragma omp parallel num_threads(nth) reduction(max:error) reduction(+:events,evals)
            {
                pair_work w{std::vector<double>(N_K), std::vector<double>(N_K), std::vector<double>(N_K),
                            std::vector<double>(N_K), std::vector<double>(N_K), {}, {}};
                std::vector<double> tmp;
                #pragma omp for schedule(dynamic, 256)
                for (int s = 0; s < count; s++)
                {
                    int i = ids[s];
                    run_owner(i, steps[spatial[i]], y, z, size, numr, active, nbr, measure, size_old.data(),
                              numr_old.data(), lambda0, &cached[4*size_t(i)], rng, seed, bath[i], w, tmp, evals);
                    error = std::max(error, bath[i].error);
                    events += bath[i].events;
                }
            }
            stats->events += events;
            stats->pair_evals += evals;
            if (error) throw std::runtime_error("local collision chain error " + std::to_string(error));
            if (dg)
            {
                // chain blocks and launches of the reference's COL_EVENT_CAP continuation scheme
                long long chained = 0, longest = 0;
                for (int i : ids)
                {
                    if (!bath[i].chained) continue;
                    chained++;
                    longest = std::max<long long>(longest, bath[i].events);
                    dg->chain_blocks += (unsigned long long)(bath[i].events / COL_EVENT_CAP + 1);
                    for (int k = 0; k < 7; k++)
                    {
                        cat_count[size_t(i)*7 + k] += bath[i].cat_count[k];
                        cat_log[size_t(i)*7 + k] += bath[i].cat_log[k];
                    }
                }
                int continuations = chained ? int(longest / COL_EVENT_CAP + 1) : 0;
                dg->chain_launches += (unsigned long long)continuations;
                dg->continuation_launches += continuations;
                dg->wave_count++;
                dg->waves++;
                dg->owner_updates += (unsigned long long)count;
            }
            // col_audit_bin (col_audit_bin.cu:13-60)
            std::fill(audit.begin(), audit.end(), audit_bin{});
            for (int i : ids)
            {
                if (!active[i]) continue;
                int g = spatial[i];
                double dur = steps[g];
                int bs = binmap[g*COL_BIN_S + size_bin(ctl, size_old[i], g)];
                int be = binmap[g*COL_BIN_S + size_bin(ctl, size[i], g)];
                double weight = numr_old[i]*grain_mass(size_old[i]);
                double probability = -std::expm1(-rate[i]*dur);
                double expected = bath[i].hazard;
                double ev = double(bath[i].events);
                audit_bin &a = audit[bs];
                a.owner_count++;
                if (!std::isfinite(weight) || !(weight > 0.0) || !std::isfinite(probability) || !std::isfinite(expected)
                    || expected < 0.0 || !std::isfinite(ev) || !std::isfinite(bath[i].j1) || bath[i].j1 < 0.0
                    || !std::isfinite(bath[i].j2) || bath[i].j2 < 0.0 || !std::isfinite(bath[i].jmax) || bath[i].jmax < 0.0
                    || !std::isfinite(size[i]) || !(size[i] > 0.0))
                { a.invalid_count++; continue; }
                a.mass += weight;
                a.predicted_f += weight*probability;
                a.predicted_e += weight*expected;
                a.predicted_var_f += weight*weight*probability*(1.0 - probability);
                a.predicted_var_e += weight*weight*expected;
                a.predicted_g += weight*bath[i].j1;
                a.predicted_var_g += weight*weight*bath[i].j2;
                a.maximum_weight = std::max(a.maximum_weight, weight);
                a.maximum_g_jump = std::max(a.maximum_g_jump, weight*bath[i].jmax);
                if (bath[i].events > 0) a.touched += weight;
                a.event_weight += weight*ev;
                a.growth += weight*std::fabs(std::log(size[i] / size_old[i]));
                audit[be].end_mass += weight;
            }
            for (int c : groups)
            {
                if (owners[c].empty()) continue;
                int first = binmap[c*COL_BIN_S], last = bin_end(c, merged);
                double mass = 0.0;
                for (int b = first; b < last; ++b)
                {
                    if (audit[b].invalid_count) throw std::runtime_error("invalid local audit state");
                    mass += audit[b].mass;
                }
                if (!(mass > 0)) continue;
                std::vector<audit_bin> slice(audit.begin() + first, audit.begin() + last);
                double before = ctl.state[c].limit_scale;
                auto result = finish_bath(slice, ctl.state[c]);
                passed[c] = !(result.activity_overshoot || result.distribution_overshoot || result.persistent_overshoot);
                stats->baths++;
                diag_group &g = dgs[c];
                g.updates++;
                g.overshoots += result.activity_overshoot || result.distribution_overshoot;
                g.persistent += result.persistent_overshoot;
                g.max_growth = std::max(g.max_growth, result.max_g);
                g.max_activity = std::max(g.max_activity, result.max_f);
                if (dg)
                    dg->baths.push_back(diag_bath{c, op_index, ++bath_index, last - first, steps[c], before,
                                                  ctl.state[c].limit_scale, result});
            }
            stats->waves++;
            stats->owner_updates += count;
            dt_col = duration;
            for (int c : groups) dt_col = std::min(dt_col, steps[c]);
            sched.advance(groups);
        }
        stats->dt_col = dt_col;
        stats->clock_dyn = duration;
        if (dg)
        {
            dg->finest_ticks = sched.end;
            for (int c = 0; c < LOCAL_GROUPS; ++c)
            {
                dgs[c].owners = int(owners[c].size());
                dgs[c].level = sched.level[c];
                dgs[c].dt = steps[c];
                dgs[c].final_request = current[c];
            }
            dg->groups = dgs;
            for (int k = 0; k < 7; k++) { dg->cat_count[k] = 0; dg->cat_log[k] = 0.0; }
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 7; k++)
                {
                    dg->cat_count[k] += cat_count[size_t(i)*7 + k];
                    dg->cat_log[k] += cat_log[size_t(i)*7 + k];
                }
        }
        return 0;
    }
    catch (const std::exception &e)
    {
        if (err && err_len > 0)
        {
            std::strncpy(err, e.what(), size_t(err_len) - 1);
            err[err_len - 1] = 0;
        }
        return 1;
    }
}

}  // extern "C"
