// Model directory "smol_const": the constant-kernel Smoluchowski benchmark (val/paper/smoluchowski/const,
// src/const_defs.cuh and src/_collision.cuh there). Unit monomers of diameter 1 with material density
// 6/pi, so grain mass = size^3; the constant synthetic kernel (COAG_KERNEL = 0); a unit-volume pair
// measure in place of the KNN measure, so every representative's summed propensity starts at one and the
// normalised number density is M0 = 1 / (1 + t/2). Particles sit on a full-azimuth vertically integrated
// annulus 0.5 <= r < 1.5 and never move. Compile-time constants keep the names of the reference's
// model header (inc/swarm/const_defs.cuh:30-159).
//
// Differences from the reference campaign: 32 neighbour slots (the campaign sweeps 16..256), 8 adaptive
// size bins per controller group (the campaign uses 64; this library's per-block controller accumulators
// hold 8), the Morton search (the campaign enforces the k-d tree; neighbour sets are the same).
#ifndef GDX_MODEL_CONST_DEFS_H
#define GDX_MODEL_CONST_DEFS_H

#define GDX_MODEL_NAME "smol_const"

#define CODE_UNIT 1
#define GDX_UNIT_MEASURE 1   // _get_ball_measure returns 1 for any owner with a neighbour (smoluchowski/const/src/_collision.cuh)

namespace gdx {

constexpr double G        = 1.0;
constexpr double M_S      = 1.0;
constexpr double R_0      = 1.0;
constexpr double S_0      = 1.0;

constexpr int    N_X      = 2;
constexpr double X_MIN    = -3.14159265358979323846;
constexpr double X_MAX    = +3.14159265358979323846;
constexpr int    N_Y      = 100;
constexpr double Y_MIN    = 0.5;
constexpr double Y_MAX    = 1.5;
constexpr int    N_Z      = 1;
constexpr double Z_MIN    = 0.5*3.14159265358979323846;
constexpr double Z_MAX    = 0.5*3.14159265358979323846;

constexpr double SIGMA_0  = 1.0e-02;
constexpr double METAL_Z  = 1.0e-02;
constexpr double ASPR_0   = 0.05;
constexpr double IDX_P    = -1.0;
constexpr double IDX_Q    = -0.4;
constexpr double ALPHA    = 1.0e-04;
constexpr double REYNOLDS_0 = 1.0e+08;

constexpr double STOKES_0 = 1.0e-03;
constexpr double RHO_0    = 6.0 / 3.14159265358979323846;   // grain mass = size^3

constexpr int    COAG_KERNEL = 0;          // constant kernel: lambda_0 * N_j / V with V = 1
constexpr int    N_K         = 32;
constexpr double H_SEARCH    = 128.0;
constexpr double V_FRAG      = 1.0;        // synthetic kernels have zero relative speed: every event sticks

constexpr int    COL_EVENT_CAP  = 32;
constexpr int    COL_BIN_X      = 8;
constexpr int    COL_BIN_Y      = 4;
constexpr int    COL_BIN_Z      = 2;
constexpr int    COL_BIN_S      = 8;
constexpr int    COL_BIN_MIN    = 64;
constexpr double COL_BATH_MAX   = 1.0e+100; // never binds: the longest bath is the operator horizon
constexpr double COL_BATH_EPS   = 0.02;
constexpr double COL_BATH_ALPHA = 1.0e-03;

constexpr int    MORTON_LEAF_TARGET = 128;
constexpr int    MORTON_MAX_LEVEL   = 20;

constexpr double INIT_SMIN = 1.0;
constexpr double INIT_SMAX = 1.0;

}  // namespace gdx

#endif  // GDX_MODEL_CONST_DEFS_H
