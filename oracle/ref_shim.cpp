// ref_shim.cpp -- C entry points around PySIT's own step kernels, so the
// oracle can be checked against the reference's code itself.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) from the
// reference headers where they lie (pysit/solvers/constant_density_acoustic/
// time/scalar/constant_density_acoustic_time_scalar_{2D,3D}.h and
// pysit/solvers/fd_tools/fd_manual.hpp); the output goes to oracle/_ref/ and is
// never committed.  The entry points take the same argument list as the
// reference's templates (constant_density_acoustic_time_scalar_2D.h:44-62,
// constant_density_acoustic_time_scalar_3D.h:47-72) minus the unused
// (rows, cols) shape pairs that SWIG passes.  ref_step2d_omp_* wraps the
// reference's OpenMP variants of orders 4 and 6
// (constant_density_acoustic_time_scalar_2D_4.h, _2D_6.h, exported by SWIG as
// constant_density_acoustic_time_scalar_2D_{4,6}omp), which take their
// thread count from OMP_NUM_THREADS (the caller must set it).
#include "constant_density_acoustic_time_scalar_2D.h"
#include "constant_density_acoustic_time_scalar_3D.h"
#include "constant_density_acoustic_time_scalar_2D_4.h"
#include "constant_density_acoustic_time_scalar_2D_6.h"

namespace {

template <typename T, int A>
void call2d(T* km1_u, T* k_phix, T* k_phiz, T* k_u, T* C, T* rhs, T* xl, int nxl,
            T* xr, int nxr, T* zl, int nzl, T* zr, int nzr, double dt, double dx,
            double dz, int nx, int nz, T* kp1_phix, T* kp1_phiz, T* kp1_u)
{
    const int n = nx * nz;
    cda_time_scalar_2D<T, A>(km1_u, n, 1, k_phix, n, 1, k_phiz, n, 1, k_u, n, 1, C,
                             n, 1, rhs, n, 1, xl, nxl, xr, nxr, zl, nzl, zr, nzr, dt,
                             dx, dz, nx, nz, kp1_phix, n, 1, kp1_phiz, n, 1, kp1_u,
                             n, 1);
}

template <typename T, int A>
void call3d(T* km1_u, T* k_phix, T* k_phiy, T* k_phiz, T* k_psi, T* k_u, T* C,
            T* rhs, T* xl, int nxl, T* xr, int nxr, T* yl, int nyl, T* yr, int nyr,
            T* zl, int nzl, T* zr, int nzr, double dt, double dx, double dy,
            double dz, int nx, int ny, int nz, T* kp1_phix, T* kp1_phiy,
            T* kp1_phiz, T* kp1_psi, T* kp1_u)
{
    const int n = nx * ny * nz;
    cda_time_scalar_3D<T, A>(km1_u, n, 1, k_phix, n, 1, k_phiy, n, 1, k_phiz, n, 1,
                             k_psi, n, 1, k_u, n, 1, C, n, 1, rhs, n, 1, xl, nxl, xr,
                             nxr, yl, nyl, yr, nyr, zl, nzl, zr, nzr, dt, dx, dy, dz,
                             nx, ny, nz, kp1_phix, n, 1, kp1_phiy, n, 1, kp1_phiz, n,
                             1, kp1_psi, n, 1, kp1_u, n, 1);
}

template <typename T>
int call2d_omp(int order, T* km1_u, T* k_phix, T* k_phiz, T* k_u, T* C, T* rhs, T* xl, int nxl,
               T* xr, int nxr, T* zl, int nzl, T* zr, int nzr, double dt, double dx,
               double dz, int nx, int nz, T* kp1_phix, T* kp1_phiz, T* kp1_u)
{
    const int n = nx * nz;
    if (order == 4)
        cda_time_scalar_2D_4<T, 4>(km1_u, n, 1, k_phix, n, 1, k_phiz, n, 1, k_u, n, 1, C, n, 1, rhs, n, 1,
                                   xl, nxl, xr, nxr, zl, nzl, zr, nzr, dt, dx, dz, nx, nz, kp1_phix, n, 1,
                                   kp1_phiz, n, 1, kp1_u, n, 1);
    else if (order == 6)
        cda_time_scalar_2D_6<T, 6>(km1_u, n, 1, k_phix, n, 1, k_phiz, n, 1, k_u, n, 1, C, n, 1, rhs, n, 1,
                                   xl, nxl, xr, nxr, zl, nzl, zr, nzr, dt, dx, dz, nx, nz, kp1_phix, n, 1,
                                   kp1_phiz, n, 1, kp1_u, n, 1);
    else
        return -1;
    return 0;
}

}  // namespace

#define REF_STEP2D_OMP(T, SUF)                                                                          \
    extern "C" int ref_step2d_omp_##SUF(int order, T* km1_u, T* k_phix, T* k_phiz, T* k_u, T* C, T* rhs, \
                                        T* xl, int nxl, T* xr, int nxr, T* zl, int nzl, T* zr, int nzr,   \
                                        double dt, double dx, double dz, int nx, int nz, T* kp1_phix,     \
                                        T* kp1_phiz, T* kp1_u)                                            \
    {                                                                                                     \
        return call2d_omp<T>(order, km1_u, k_phix, k_phiz, k_u, C, rhs, xl, nxl, xr, nxr, zl, nzl, zr,   \
                             nzr, dt, dx, dz, nx, nz, kp1_phix, kp1_phiz, kp1_u);                         \
    }

#define REF_STEP2D(T, SUF)                                                           \
    extern "C" int ref_step2d_##SUF(int order, T* km1_u, T* k_phix, T* k_phiz,       \
                                    T* k_u, T* C, T* rhs, T* xl, int nxl, T* xr,     \
                                    int nxr, T* zl, int nzl, T* zr, int nzr,         \
                                    double dt, double dx, double dz, int nx, int nz, \
                                    T* kp1_phix, T* kp1_phiz, T* kp1_u)              \
    {                                                                                \
        switch (order) {                                                             \
        case 2: call2d<T, 2>(km1_u, k_phix, k_phiz, k_u, C, rhs, xl, nxl, xr, nxr,   \
                             zl, nzl, zr, nzr, dt, dx, dz, nx, nz, kp1_phix,         \
                             kp1_phiz, kp1_u); return 0;                             \
        case 4: call2d<T, 4>(km1_u, k_phix, k_phiz, k_u, C, rhs, xl, nxl, xr, nxr,   \
                             zl, nzl, zr, nzr, dt, dx, dz, nx, nz, kp1_phix,         \
                             kp1_phiz, kp1_u); return 0;                             \
        case 6: call2d<T, 6>(km1_u, k_phix, k_phiz, k_u, C, rhs, xl, nxl, xr, nxr,   \
                             zl, nzl, zr, nzr, dt, dx, dz, nx, nz, kp1_phix,         \
                             kp1_phiz, kp1_u); return 0;                             \
        case 8: call2d<T, 8>(km1_u, k_phix, k_phiz, k_u, C, rhs, xl, nxl, xr, nxr,   \
                             zl, nzl, zr, nzr, dt, dx, dz, nx, nz, kp1_phix,         \
                             kp1_phiz, kp1_u); return 0;                             \
        default: return -1;                                                          \
        }                                                                            \
    }

#define REF_STEP3D(T, SUF)                                                           \
    extern "C" int ref_step3d_##SUF(                                                 \
        int order, T* km1_u, T* k_phix, T* k_phiy, T* k_phiz, T* k_psi, T* k_u,      \
        T* C, T* rhs, T* xl, int nxl, T* xr, int nxr, T* yl, int nyl, T* yr,         \
        int nyr, T* zl, int nzl, T* zr, int nzr, double dt, double dx, double dy,    \
        double dz, int nx, int ny, int nz, T* kp1_phix, T* kp1_phiy, T* kp1_phiz,    \
        T* kp1_psi, T* kp1_u)                                                        \
    {                                                                                \
        switch (order) {                                                             \
        case 2: call3d<T, 2>(km1_u, k_phix, k_phiy, k_phiz, k_psi, k_u, C, rhs, xl,  \
                             nxl, xr, nxr, yl, nyl, yr, nyr, zl, nzl, zr, nzr, dt,   \
                             dx, dy, dz, nx, ny, nz, kp1_phix, kp1_phiy, kp1_phiz,   \
                             kp1_psi, kp1_u); return 0;                              \
        case 4: call3d<T, 4>(km1_u, k_phix, k_phiy, k_phiz, k_psi, k_u, C, rhs, xl,  \
                             nxl, xr, nxr, yl, nyl, yr, nyr, zl, nzl, zr, nzr, dt,   \
                             dx, dy, dz, nx, ny, nz, kp1_phix, kp1_phiy, kp1_phiz,   \
                             kp1_psi, kp1_u); return 0;                              \
        case 6: call3d<T, 6>(km1_u, k_phix, k_phiy, k_phiz, k_psi, k_u, C, rhs, xl,  \
                             nxl, xr, nxr, yl, nyl, yr, nyr, zl, nzl, zr, nzr, dt,   \
                             dx, dy, dz, nx, ny, nz, kp1_phix, kp1_phiy, kp1_phiz,   \
                             kp1_psi, kp1_u); return 0;                              \
        case 8: call3d<T, 8>(km1_u, k_phix, k_phiy, k_phiz, k_psi, k_u, C, rhs, xl,  \
                             nxl, xr, nxr, yl, nyl, yr, nyr, zl, nzl, zr, nzr, dt,   \
                             dx, dy, dz, nx, ny, nz, kp1_phix, kp1_phiy, kp1_phiz,   \
                             kp1_psi, kp1_u); return 0;                              \
        default: return -1;                                                          \
        }                                                                            \
    }

REF_STEP2D(double, d)
REF_STEP2D(float, f)
REF_STEP2D_OMP(double, d)
REF_STEP2D_OMP(float, f)
REF_STEP3D(double, d)
REF_STEP3D(float, f)
