// Model directory "ensemble_2d": the plain_2d disk with 32 neighbour slots, for ensembles of many
// realizations of that distribution. Vertically integrated radial-azimuthal disk (N_Z = 1), constant
// synthetic kernel (COAG_KERNEL = 0), coagulation only. Compile-time constants
// keep the names of the reference's model header (inc/swarm/const_defs.cuh:30-159).
//
// Workload: realizations of floor(2^U[14, 20]) representatives each, r uniform on [0.5, 2.0), phi uniform,
// monodisperse sizes; each realization is its own swarm (its own lambda_0 = N / N_K / M).
#ifndef GDX_MODEL_CONST_DEFS_H
#define GDX_MODEL_CONST_DEFS_H

#define GDX_MODEL_NAME "ensemble_2d"

#define CODE_UNIT 1

namespace gdx {

constexpr double G        = 1.0;
constexpr double M_S      = 1.0;
constexpr double R_0      = 1.0;
constexpr double S_0      = 1.0;

constexpr int    N_X      = 100;
constexpr double X_MIN    = -3.14159265358979323846;
constexpr double X_MAX    = +3.14159265358979323846;
constexpr int    N_Y      = 100;
constexpr double Y_MIN    = 0.5;
constexpr double Y_MAX    = 2.0;
constexpr int    N_Z      = 1;
constexpr double Z_MIN    = 0.5*3.14159265358979323846;
constexpr double Z_MAX    = 0.5*3.14159265358979323846;

constexpr double SIGMA_0  = 1.0e-02;
constexpr double METAL_Z  = 1.0e-02;
constexpr double ASPR_0   = 0.05;
constexpr double IDX_P    = -1.0;
constexpr double IDX_Q    = -0.5;
constexpr double ALPHA    = 1.0e-03;
constexpr double REYNOLDS_0 = 1.0e+08;

constexpr double STOKES_0 = 1.0e-03;
constexpr double RHO_0    = 1.0;

constexpr int    COAG_KERNEL = 0;          // constant kernel: lambda_0 * N_j / V
constexpr int    N_K         = 32;
constexpr double H_SEARCH    = 1.0;
constexpr double V_FRAG      = 1.0e+300;   // coagulation only

constexpr int    COL_EVENT_CAP  = 32;
constexpr int    COL_BIN_X      = 8;
constexpr int    COL_BIN_Y      = 4;
constexpr int    COL_BIN_Z      = 2;
constexpr int    COL_BIN_S      = 8;
constexpr int    COL_BIN_MIN    = 64;
constexpr double COL_BATH_MAX   = 0.05;
constexpr double COL_BATH_EPS   = 0.02;
constexpr double COL_BATH_ALPHA = 1.0e-03;

constexpr int    MORTON_LEAF_TARGET = 128;
constexpr int    MORTON_MAX_LEVEL   = 20;

constexpr double INIT_SMIN = 1.0e-04;
constexpr double INIT_SMAX = 1.0e-04;

}  // namespace gdx

#endif  // GDX_MODEL_CONST_DEFS_H
