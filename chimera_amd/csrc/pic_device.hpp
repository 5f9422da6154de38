// pic_device.hpp -- per-particle device arithmetic shared by the drop-in
// particle kernels and the fused kernels of resident simulations.
#pragma once

#include <cstring>

#include "common.hpp"
#include "../../include/chimera_hip.h"

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

// ---------------------------------------------------------------- external-field devices
//
// The devices of devices.f90 evaluated at a particle, each adding its field to
// f (Ex, Ey, Ez, Bx, By, Bz) the way the reference's subroutine adds it to
// Fld; the parameters are prepared on the host (make_ext_device below).
enum { DEV_UNDUL = 0, DEV_UNDUL_TAPER = 1, DEV_MAPPED = 2, DEV_MAPPED_TAP = 3, DEV_PLANEWAVE = 4,
       DEV_GAUSSBEAM = 5 };
struct ExtDevice {
    int kind, nmap;
    const double* map;    // DEV_MAPPED*: a0(2, nmap) in device memory
    double p[10];
};
constexpr int MAX_EXT_DEVICES = 4;
struct ExtDevices {
    ExtDevice d[MAX_EXT_DEVICES];
    int n;
    double t;             // time of the devices that depend on it (i_step * TimeStep)
};

// devices.f90:162-203 (analytic) and :117-160 (tapered): p = a0, lambda, X0, Lx, ku, taper
__device__ __forceinline__ void ext_undul(const double* p, bool tapered, double x, double y, double df[6]) {
    const double a0 = p[0], lambda = p[1], X0 = p[2], Lx = p[3], ku = p[4];
    double ampl;
    if (x <= X0 || x >= X0 + Lx) ampl = 0.0;
    else if (x > X0 && x < X0 + lambda) ampl = (x - X0) / lambda;
    else if (x > X0 + Lx - lambda && x < X0 + Lx) ampl = (X0 + Lx - x) / lambda;
    else ampl = 1.0;
    if (tapered) ampl = ampl * (1.0 + p[5] * (x - X0 - 0.5 * Lx) / (0.5 * Lx));
    ampl = ampl * a0;
    df[4] = ampl * sin(ku * (x - X0)) * cosh(ku * y);
    df[3] = ampl * cos(ku * (x - X0)) * sinh(ku * y);
}

// devices.f90:18-62 (tap = false) and :64-115 (tap = true): the field map a0(2, nx)
// on nodes Xleft + (i-1) dx, i = 1..nx, with quadratic weights around node ix;
// p = Xleft, dx, dx_inv, Xright, ku, Lx, taper, x_shift.  Particles outside
// [Xleft + dx, Xright - dx] see no field; a node outside 1..nx contributes zero.
__device__ __forceinline__ void ext_mapped(const double* p, const double* __restrict__ a0, int nx, bool tap, double xp,
                                           double yp, double df[6]) {
    const double Xleft = p[0], dx = p[1], dx_inv = p[2], Xright = p[3], ku = p[4];
    if (xp < Xleft + dx || xp > Xright - dx) return;
    const double fx = floor((xp - Xleft) * dx_inv + 0.5);
    if (!(fx >= 0.0 && fx <= (double)nx + 1.0)) return;       // NaN or off the map: no node in range
    const int ix = (int)fx;
    const double ddx = (xp - Xleft) * dx_inv - ix;
    const double s[3] = {0.5 * ((0.5 - ddx) * (0.5 - ddx)), 0.75 - ddx * ddx, 0.5 * ((0.5 + ddx) * (0.5 + ddx))};
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int j = ix - 1 + k;                                // 1-based node
        const double v1 = (j >= 1 && j <= nx) ? a0[2 * (int64_t)(j - 1)] : 0.0;
        const double v2 = (j >= 1 && j <= nx) ? a0[2 * (int64_t)(j - 1) + 1] : 0.0;
        b1 = b1 + s[k] * v1;
        b2 = b2 + s[k] * v2;
    }
    if (tap) {
        const double Lx = p[5], taper = p[6], x_shift = p[7];
        const double amp = 1.0 + taper * (xp - x_shift - Xleft - 0.5 * Lx) / (0.5 * Lx);
        df[4] = amp * b1 * cosh(ku * yp);
        df[3] = amp * b2 * sinh(ku * yp);
    } else {
        df[4] = b1 * cosh(ku * yp);
        df[3] = b2 * sinh(ku * yp);
    }
}

// devices.f90:205-252: p = a0, X0, Lx, ramp, sin(theta), cos(theta), k0, phi0
__device__ __forceinline__ void ext_planewave(const double* p, double t, double x, double y, double df[6]) {
    const double a0 = p[0], X0 = p[1], Lx = p[2], ramp = p[3], sinth = p[4], costh = p[5], k0 = p[6], phi0 = p[7];
    double ampl;
    if (x <= X0 || x >= X0 + Lx) ampl = 0.0;
    else if (x > X0 && x < X0 + ramp) ampl = (x - X0) / ramp;
    else if (x > X0 + Lx - ramp && x < X0 + Lx) ampl = (X0 + Lx - x) / ramp;
    else ampl = 1.0;
    ampl = a0 * ampl * sin(k0 * (x * costh + y * sinth - t) + phi0);
    df[2] = ampl;
    df[3] = ampl * sinth;
    df[4] = -(ampl * costh);
}

// devices.f90:254-297: p = a0, lambda, axis, x0, y0, z0, 1/Lx^2, 1/Ly^2, 1/Lz^2
__device__ __forceinline__ void ext_gaussbeam(const double* p, double t, double x, double y, double z, double df[6]) {
    const double a0 = p[0], lambda = p[1], axis = p[2];
    const double xp = x - p[3] - axis * t, yp = y - p[4], zp = z - p[5];
    const double e = a0 * exp(-(xp * xp * p[6]) - yp * yp * p[7] - zp * zp * p[8]) * sin(2.0 * M_PI / lambda * xp);
    df[2] = e;
    df[4] = -(axis * e);
}

// One device's contribution at (x, y, z); df is zeroed here, so f + df equals
// the reference's f = f + term for every component a device writes.
__device__ __forceinline__ void ext_device_field(const ExtDevice& d, double t, double x, double y, double z,
                                                 double df[6]) {
#pragma unroll
    for (int c = 0; c < 6; c++) df[c] = 0.0;
    switch (d.kind) {
        case DEV_UNDUL: ext_undul(d.p, false, x, y, df); break;
        case DEV_UNDUL_TAPER: ext_undul(d.p, true, x, y, df); break;
        case DEV_MAPPED: ext_mapped(d.p, d.map, d.nmap, false, x, y, df); break;
        case DEV_MAPPED_TAP: ext_mapped(d.p, d.map, d.nmap, true, x, y, df); break;
        case DEV_PLANEWAVE: ext_planewave(d.p, t, x, y, df); break;
        default: ext_gaussbeam(d.p, t, x, y, z, df); break;
    }
}

// Every device of D in order, added to the particle's field f (FP64 terms,
// added in the precision of f).
template <typename R>
__device__ __forceinline__ void ext_fields(const ExtDevices& D, double x, double y, double z, R f[6]) {
    for (int k = 0; k < D.n; k++) {
        double df[6];
        ext_device_field(D.d[k], D.t, x, y, z, df);
        switch (D.d[k].kind) {
            case DEV_PLANEWAVE:
                f[2] += (R)df[2];
                f[3] += (R)df[3];
                f[4] += (R)df[4];
                break;
            case DEV_GAUSSBEAM:
                f[2] += (R)df[2];
                f[4] += (R)df[4];
                break;
            default:
                f[4] += (R)df[4];
                f[3] += (R)df[3];
                break;
        }
    }
}

// Host: one device from its kind, the reference's parameter array and, for the
// mapped undulators, the device copy of the map (a0 rows of nmap nodes); the
// Gaussian beam's scalar amplitude is passed as a0.  EINVAL on a short array.
int32_t make_ext_device(ExtDevice& d, int kind, const double* params, int nparams, const double* map, int nmap,
                        double a0);

// Geometry of a resident launch: Rgrid in device memory, its first and last
// node given on the host.
inline GridGeom resident_geom(double leftX, const double* Rgrid, double dx_inv, double dr_inv, double r0, double rmax,
                              int nx, int nr, int nkO) {
    GridGeom g = make_geom(leftX, Rgrid, dx_inv, dr_inv, nx, nr, nkO);
    g.r0 = r0;
    g.rmax = rmax;
    return g;
}

// The device list of a resident gather-push from the C-ABI records (host
// memory; a mapped undulator's map in device memory), at time t.
inline int32_t ext_devices(ExtDevices& devs, const chimera_device_t* d, int32_t ndev, double t) {
    memset(&devs, 0, sizeof devs);
    if (ndev < 0 || ndev > MAX_EXT_DEVICES || (ndev > 0 && d == nullptr))
        return fail(ST_EINVAL, "gather_push: %d devices (at most %d per species)", ndev, MAX_EXT_DEVICES);
    for (int k = 0; k < ndev; k++) {
        if ((d[k].kind == DEV_MAPPED || d[k].kind == DEV_MAPPED_TAP) && !is_device_ptr(d[k].map))
            return fail(ST_EINVAL, "gather_push: the map of a mapped undulator is device memory");
        CH_TRY(make_ext_device(devs.d[k], d[k].kind, d[k].params, 8, d[k].map, d[k].nmap, d[k].a0));
    }
    devs.n = ndev;
    devs.t = t;
    return ST_OK;
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
