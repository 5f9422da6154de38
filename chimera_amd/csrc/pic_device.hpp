// pic_device.hpp -- per-particle device arithmetic shared by the drop-in
// particle kernels and the fused kernels of resident simulations.
#pragma once

#include "common.hpp"

namespace chimera {

constexpr int MAX_MODES = 8;  // azimuthal modes m = 0..7

// Element (particle ip, component c) of a particle array sits at ip*p + c*c_.
struct PStride {
    int64_t p, c;
    __host__ __device__ int64_t at(int64_t ip, int comp) const { return ip * p + comp * c; }
};

// The quasi-cylindrical grid as grid_deps.f90 sees it: x nodes 0..nx, radial
// nodes 0..nr at Rgrid (Rgrid(0) = -dr/2 is the ghost), modes 0..nkO.
struct GridGeom {
    double leftX, dx_inv, dr_inv;
    double r0, rmax;  // Rgrid(0), Rgrid(nr)
    const double* Rgrid;
    // When set, the resident kernels take leftX from this device address at
    // run time instead (a moving window advanced on the device, so that a
    // captured step replays correctly); the drop-in kernels ignore it.
    const double* leftX_dev;
    int nx, nr, nkO;
};

inline GridGeom make_geom(double leftX, const double* Rgrid_dev, double dx_inv, double dr_inv, int nx, int nr, int nkO) {
    GridGeom g;
    g.leftX = leftX;
    g.dx_inv = dx_inv;
    g.dr_inv = dr_inv;
    g.Rgrid = Rgrid_dev;
    g.leftX_dev = nullptr;
    g.r0 = 0.0;
    g.rmax = 0.0;
    g.nx = nx;
    g.nr = nr;
    g.nkO = nkO;
    return g;
}

// Fill r0 / rmax from the caller's Rgrid (host or device memory).
inline int32_t load_rgrid_edges(GridGeom& g, const double* Rgrid_caller, const double* Rgrid_dev) {
    if (is_device_ptr(Rgrid_caller)) {
        CH_HIP(hipMemcpy(&g.r0, Rgrid_dev, sizeof(double), hipMemcpyDeviceToHost));
        CH_HIP(hipMemcpy(&g.rmax, Rgrid_dev + g.nr, sizeof(double), hipMemcpyDeviceToHost));
    } else {
        g.r0 = Rgrid_caller[0];
        g.rmax = Rgrid_caller[g.nr];
    }
    return ST_OK;
}

struct PartGeom {
    int64_t ix, ir;
    double S[2][2];             // S[0][a]: x weight of node ix+a; S[1][b]: r weight of node ir+b
    double2 ph[MAX_MODES];      // ((y + sgn*i*z)/r)^m, phase 0 at r = 0
};

// The window's left edge a resident kernel uses: the device value when there is one.
__device__ __forceinline__ GridGeom frame_geom(GridGeom g) {
    if (g.leftX_dev) g.leftX = *g.leftX_dev;
    return g;
}

// Cell indices, linear weights and azimuthal phases of one particle
// (grid_deps.f90:41-66 and :175-199).  Returns false for particles the
// reference skips (r >= Rgrid(nr)) and for particles whose cell lies outside
// the grid arrays (the reference has no x bounds check and would write or
// read outside the array there); NaN coordinates fall in the second group.
__device__ __forceinline__ bool part_geom(PartGeom& pg, double xp, double yp, double zp, const GridGeom& g,
                                          double sgn) {
    const double rp = sqrt(yp * yp + zp * zp);
    if (!(rp < g.rmax)) return false;
    const double fx = (xp - g.leftX) * g.dx_inv;
    const double fr = (rp - g.r0) * g.dr_inv;
    if (!(fx >= 0.0 && fx < (double)g.nx)) return false;
    if (!(fr >= 0.0 && fr < (double)g.nr)) return false;
    const double fix = floor(fx), fir = floor(fr);
    pg.ix = (int64_t)fix;
    pg.ir = (int64_t)fir;
    pg.S[0][1] = fx - fix;
    pg.S[0][0] = 1.0 - pg.S[0][1];
    pg.S[1][1] = (rp - g.Rgrid[pg.ir]) * g.dr_inv;
    pg.S[1][0] = 1.0 - pg.S[1][1];
    double2 p1 = rp > 0.0 ? cmk(yp / rp, sgn * zp / rp) : cmk(0.0, 0.0);
    pg.ph[0] = cmk(1.0, 0.0);
#pragma unroll
    for (int m = 1; m < MAX_MODES; m++)
        if (m <= g.nkO) pg.ph[m] = cmul(pg.ph[m - 1], p1);
    return true;
}

// Relativistic Boris rotation with half-step dt_2 = dt/2 (particle_tools.f90:37-51),
// in the precision R of the momenta.
template <typename R>
__device__ __forceinline__ void boris(R u[3], const R f[6], R dt_2) {
    R Um[3], t[3], s[3], U0[3];
    for (int c = 0; c < 3; c++) Um[c] = u[c] + dt_2 * f[c];
    const R gamma = sqrt((R)1 + (Um[0] * Um[0] + Um[1] * Um[1] + Um[2] * Um[2]));
    for (int c = 0; c < 3; c++) t[c] = dt_2 * f[3 + c] / gamma;
    const R t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    for (int c = 0; c < 3; c++) s[c] = 2 * t[c] / (1 + t2);
    U0[0] = Um[0] + Um[1] * t[2] - Um[2] * t[1];
    U0[1] = Um[1] - Um[0] * t[2] + Um[2] * t[0];
    U0[2] = Um[2] + Um[0] * t[1] - Um[1] * t[0];
    const R Up0 = Um[0] + U0[1] * s[2] - U0[2] * s[1];
    const R Up1 = Um[1] - U0[0] * s[2] + U0[2] * s[0];
    const R Up2 = Um[2] + U0[0] * s[1] - U0[1] * s[0];
    u[0] = Up0 + dt_2 * f[0];
    u[1] = Up1 + dt_2 * f[1];
    u[2] = Up2 + dt_2 * f[2];
}

// Workgroups of a 1-D launch go to the 8 XCDs round-robin (workgroup b to XCD
// b % 8), and each XCD has its own L2.  xcd_block renumbers them so that XCD x
// takes one contiguous range of the logical block indices: with particles in
// cell order that is one region of the grid per L2.  A bijection of [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

struct UndulParams { double a0, lambda, X0, Lx, ku, taper; int tapered; };

inline UndulParams make_undul(const double p[4]) {
    UndulParams u;
    u.a0 = p[0];
    u.lambda = p[1];
    u.X0 = p[2];
    u.Lx = p[3];
    u.ku = 2.0 * M_PI / p[1];
    u.taper = 0.0;
    u.tapered = 0;
    return u;
}

// params = (a0, lambda, X0, Lx, taper) of the linearly tapered undulator.
inline UndulParams make_undul_taper(const double p[5]) {
    UndulParams u = make_undul(p);
    u.taper = p[4];
    u.tapered = 1;
    return u;
}

// Analytic planar undulator (devices.f90:162-203), evaluated in FP64 and added
// to the gathered fields; tapered, the amplitude is scaled by
// 1 + taper (x - X0 - Lx/2) / (Lx/2) (devices.f90:117-160).
template <typename R>
__device__ __forceinline__ void undul_field(const UndulParams& up, double x, double y, R f[6]) {
    double ampl;
    if (x <= up.X0 || x >= up.X0 + up.Lx) ampl = 0.0;
    else if (x > up.X0 && x < up.X0 + up.lambda) ampl = (x - up.X0) / up.lambda;
    else if (x > up.X0 + up.Lx - up.lambda && x < up.X0 + up.Lx) ampl = (up.X0 + up.Lx - x) / up.lambda;
    else ampl = 1.0;
    if (up.tapered) ampl = ampl * (1.0 + up.taper * (x - up.X0 - 0.5 * up.Lx) / (0.5 * up.Lx));
    ampl = ampl * up.a0;
    f[4] += (R)(ampl * sin(up.ku * (x - up.X0)) * cosh(up.ku * y));
    f[3] += (R)(ampl * cos(up.ku * (x - up.X0)) * sinh(up.ku * y));
}

// Launchers of the particle kernels (particles.hip).
int32_t launch_push_coords(double* x, PStride sx, const double* u, PStride su, double* xc, PStride sc, double dt,
                           int np);
int32_t launch_dep_curr(const double* x, PStride sx, const double* u, PStride su, const double* w, double2* curr,
                        const GridGeom& g, int ip0, int ip1);
template <typename W>
int32_t launch_dep_dens(const double* x, PStride sx, const W* w, double2* dens, const GridGeom& g, int ip0, int ip1);
// a[ix, ir, line] *= v[ir] on nlines (nx1, nr1) planes
int32_t scale_rows(double2* a, const double* v, int nx1, int nr1, int nlines);

}  // namespace chimera
