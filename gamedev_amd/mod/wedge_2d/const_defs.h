// Model directory "wedge_2d": vertically integrated azimuthal wedge x in [-0.1, 0.1] with periodic
// seams, physical (Ormel-Cuzzi) coagulation kernel, 32 neighbour slots. The disk and kernel constants
// follow the reference's wedge chain test (val/swarm/src/const_defs.cuh:38-216 under
// TEST_COLCHAIN_WEDGE_2D); the neighbour count is 32 instead of that test's 200. Compile-time
// constants keep the names of the reference's model header (inc/swarm/const_defs.cuh:30-159).
//
// The wedge is narrow: two images of one particle can share a search ball (2*0.5*sin(0.1) < 2*max cap),
// so the neighbour search keeps 3*N_K candidates and removes duplicate images.
#ifndef GDX_MODEL_CONST_DEFS_H
#define GDX_MODEL_CONST_DEFS_H

#define GDX_MODEL_NAME "wedge_2d"

#define CODE_UNIT 1

namespace gdx {

constexpr double G        = 1.0;
constexpr double M_S      = 1.0;
constexpr double R_0      = 1.0;
constexpr double S_0      = 1.0;

constexpr int    N_X      = 32;
constexpr double X_MIN    = -0.1;
constexpr double X_MAX    = +0.1;
constexpr int    N_Y      = 32;
constexpr double Y_MIN    = 0.5;
constexpr double Y_MAX    = 1.5;
constexpr int    N_Z      = 1;
constexpr double Z_MIN    = 0.5*3.14159265358979323846;
constexpr double Z_MAX    = 0.5*3.14159265358979323846;

constexpr double SIGMA_0  = 1.0;
constexpr double METAL_Z  = 1.0e-02;
constexpr double ASPR_0   = 0.05;
constexpr double IDX_P    = 2.0;
constexpr double IDX_Q    = -1.0;
constexpr double ALPHA    = 1.0e-04;
constexpr double REYNOLDS_0 = 1.0e+08;

constexpr double STOKES_0 = 0.2;
constexpr double RHO_0    = 1.0;

constexpr int    COAG_KERNEL = 3;          // physical kernel
constexpr int    N_K         = 32;
constexpr double H_SEARCH    = 1.0;
constexpr double V_FRAG      = 1.0;

constexpr int    COL_EVENT_CAP  = 32;
constexpr int    COL_BIN_X      = 8;
constexpr int    COL_BIN_Y      = 4;
constexpr int    COL_BIN_Z      = 2;
constexpr int    COL_BIN_S      = 8;
constexpr int    COL_BIN_MIN    = 64;
constexpr double COL_BATH_MAX   = 0.05;
constexpr double COL_BATH_EPS   = 0.06;
constexpr double COL_BATH_ALPHA = 1.0e-03;

constexpr int    MORTON_LEAF_TARGET = 128;
constexpr int    MORTON_MAX_LEVEL   = 20;

constexpr double INIT_SMIN = 0.05;
constexpr double INIT_SMAX = 6.4;

}  // namespace gdx

#endif  // GDX_MODEL_CONST_DEFS_H
