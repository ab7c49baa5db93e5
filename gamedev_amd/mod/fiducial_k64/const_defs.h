// Model directory "fiducial_k64": the fiducial 3D disk (physical Ormel-Cuzzi kernel with
// fragmentation) with 64 neighbour slots, one owner per wavefront in the rate and chain kernels.
// Compile-time constants keep the names of the reference's model header
// (inc/swarm/const_defs.cuh:30-159); a model directory replaces this file as a whole.
#ifndef GDX_MODEL_CONST_DEFS_H
#define GDX_MODEL_CONST_DEFS_H

#define GDX_MODEL_NAME "fiducial_k64"

// reference flags.mk switches that change the collision path
#define CODE_UNIT 1          // Reynolds number from REYNOLDS_0; no Brownian term (inc/swarm/_collision.cuh:136-150, 318-322)

namespace gdx {

constexpr double G        = 1.0;
constexpr double M_S      = 1.0;
constexpr double R_0      = 1.0;
constexpr double S_0      = 1.0e-04;     // reference size: St spans ~1e-4..1 over the size range

constexpr int    N_X      = 100;
constexpr double X_MIN    = -3.14159265358979323846;
constexpr double X_MAX    = +3.14159265358979323846;
constexpr int    N_Y      = 100;
constexpr double Y_MIN    = 0.3;
constexpr double Y_MAX    = 3.0;
constexpr int    N_Z      = 100;
constexpr double Z_MIN    = 0.5*3.14159265358979323846 - 0.4;
constexpr double Z_MAX    = 0.5*3.14159265358979323846 + 0.4;

constexpr double SIGMA_0  = 1.0e-02;
constexpr double METAL_Z  = 1.0e-02;
constexpr double ASPR_0   = 0.05;
constexpr double IDX_P    = -1.0;
constexpr double IDX_Q    = -0.5;         // h/r = ASPR_0 (R/R_0)^0.25
constexpr double ALPHA    = 1.0e-03;
constexpr double REYNOLDS_0 = 1.0e+08;

constexpr double STOKES_0 = 1.0e-03;
constexpr double RHO_0    = 1.0;

constexpr int    COAG_KERNEL = 3;          // physical kernel
constexpr int    N_K         = 64;
constexpr double H_SEARCH    = 1.0;
constexpr double V_FRAG      = 1.0e-03;    // fragmentation threshold in units of v_K(R_0) = 1

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

constexpr double INIT_SMIN = 1.0e-05;      // monomer floor used by catastrophic fragmentation
constexpr double INIT_SMAX = 1.0e-01;

}  // namespace gdx

#endif  // GDX_MODEL_CONST_DEFS_H
