// particles.hip -- particle kernels behind the f2py entry points of
// particle_tools.f90, grid_deps.f90 and devices.f90: Boris and leap-frog
// pushes, field gather, charge/current deposition, the mode normalisation of
// the gathered field, the analytic undulator and the FixedCell lattice.
//
// Particle arrays are addressed through (element stride, component stride)
// pairs, so the same kernels serve the reference's (3,np)/(6,np) Fortran
// layout (strides ncomp, 1) and the structure-of-arrays layout of resident
// simulations (strides 1, ld).
#include "common.hpp"
#include "fb_kernels.hpp"
#include "pic_device.hpp"
#include "../../include/chimera_hip.h"

namespace chimera {

// ----------------------------------------------------------------- pushes

// particle_tools.f90:18-56
__global__ void k_push_velocs(double* __restrict__ u, PStride su, const double* __restrict__ F, PStride sf,
                              double dt, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    double m[3], f[6];
    for (int c = 0; c < 3; c++) m[c] = u[su.at(ip, c)];
    for (int c = 0; c < 6; c++) f[c] = F[sf.at(ip, c)];
    boris(m, f, 0.5 * dt);
    for (int c = 0; c < 3; c++) u[su.at(ip, c)] = m[c];
}

// particle_tools.f90:58-82
__global__ void k_push_coords(double* __restrict__ x, PStride sx, const double* __restrict__ u, PStride su,
                              double* __restrict__ xc, PStride sc, double dt, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    const double u0 = u[su.at(ip, 0)], u1 = u[su.at(ip, 1)], u2 = u[su.at(ip, 2)];
    const double dt_gp = dt / sqrt(1.0 + (u0 * u0 + u1 * u1 + u2 * u2));
    const double uu[3] = {u0, u1, u2};
    for (int c = 0; c < 3; c++) {
        const double x0 = x[sx.at(ip, c)];
        const double x1 = x0 + uu[c] * dt_gp;
        x[sx.at(ip, c)] = x1;
        xc[sc.at(ip, c)] = 0.5 * (x0 + x1);
    }
}

// ----------------------------------------------------------------- gather

// grid_deps.f90:149-217.  Fld(0:nx,0:nr,0:nkO,6) complex; Fld_tot += gathered.
__global__ void k_proj_fld(const double* __restrict__ coord, PStride sx, const double* __restrict__ w,
                           const double2* __restrict__ Fld, double* __restrict__ Ftot, PStride sf, GridGeom g,
                           int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    const double wp = w[ip];
    if (wp == 0.0) return;
    PartGeom pg;
    if (!part_geom(pg, coord[sx.at(ip, 0)], coord[sx.at(ip, 1)], coord[sx.at(ip, 2)], g, +1.0)) return;
    const int64_t sr = g.nx + 1, sm = sr * (g.nr + 1), sl = sm * (g.nkO + 1);
    const double2* base = Fld + pg.ix + pg.ir * sr;
    for (int l = 0; l < 6; l++) {
        double acc = 0.0;
        for (int iO = 0; iO <= g.nkO; iO++) {
            const double2* f = base + iO * sm + l * sl;
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) {
                    const double s = pg.S[1][b] * pg.S[0][a];
                    const double2 pc = crs(s, pg.ph[iO]);
                    const double2 v = f[a + b * sr];
                    acc += pc.x * v.x - pc.y * v.y;
                }
        }
        Ftot[sf.at(ip, l)] += acc;
    }
}

// ----------------------------------------------------------------- deposition

// grid_deps.f90:18-87 (particle loop); sums with hardware FP64 atomics.
__global__ void k_dep_curr(const double* __restrict__ coord, PStride sx, const double* __restrict__ u, PStride su,
                           const double* __restrict__ w, double2* __restrict__ curr, GridGeom g, int ip0, int ip1) {
    const int ip = ip0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= ip1) return;
    const double wp = w[ip];
    if (wp == 0.0) return;
    double v[3] = {u[su.at(ip, 0)], u[su.at(ip, 1)], u[su.at(ip, 2)]};
    if ((fabs(v[0]) + fabs(v[1]) + fabs(v[2])) == 0.0) return;
    PartGeom pg;
    if (!part_geom(pg, coord[sx.at(ip, 0)], coord[sx.at(ip, 1)], coord[sx.at(ip, 2)], g, -1.0)) return;
    const double gp = sqrt(1 + v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int c = 0; c < 3; c++) v[c] = v[c] * wp / gp;
    const int64_t sr = g.nx + 1, sm = sr * (g.nr + 1), sl = sm * (g.nkO + 1);
    double* base = reinterpret_cast<double*>(curr + pg.ix + pg.ir * sr);
    for (int l = 0; l < 3; l++)
        for (int iO = 0; iO <= g.nkO; iO++) {
            const double2 pv = crs(v[l], pg.ph[iO]);
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) {
                    const double2 d = crs(pg.S[0][a] * pg.S[1][b], pv);
                    double* dst = base + 2 * (a + b * sr + iO * sm + l * sl);
                    unsafeAtomicAdd(dst, d.x);
                    unsafeAtomicAdd(dst + 1, d.y);
                }
        }
}

// grid_deps.f90:89-147 (particle loop)
template <typename W>
__global__ void k_dep_dens(const double* __restrict__ coord, PStride sx, const W* __restrict__ w,
                           double2* __restrict__ dens, GridGeom g, int ip0, int ip1) {
    const int ip = ip0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= ip1) return;
    const double wp = (double)w[ip];
    if (wp == 0.0) return;
    PartGeom pg;
    if (!part_geom(pg, coord[sx.at(ip, 0)], coord[sx.at(ip, 1)], coord[sx.at(ip, 2)], g, -1.0)) return;
    const int64_t sr = g.nx + 1, sm = sr * (g.nr + 1);
    double* base = reinterpret_cast<double*>(dens + pg.ix + pg.ir * sr);
    for (int iO = 0; iO <= g.nkO; iO++)
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) {
                const double2 d = crs(pg.S[0][a] * pg.S[1][b] * wp, pg.ph[iO]);
                double* dst = base + 2 * (a + b * sr + iO * sm);
                unsafeAtomicAdd(dst, d.x);
                unsafeAtomicAdd(dst + 1, d.y);
            }
}

// Ghost fold after deposition: row 1 -= row 0, row 0 = 0.  grid_deps.f90:80-85, :144-145.
template <typename C>
__global__ void k_ghost_fold(C* __restrict__ a, int nx1, int nr1, int nlines) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx1 * nlines) return;
    const int64_t ix = t % nx1, q = t / nx1;
    C* r0 = a + ix + q * (int64_t)nx1 * nr1;
    r0[nx1] = csub(r0[nx1], r0[0]);
    r0[0].x = 0;
    r0[0].y = 0;
}

template <typename R>
int32_t ghost_fold(cplx_t<R>* a, int nx1, int nr1, int64_t nlines) {
    if (nlines <= 0) return ST_OK;
    hipLaunchKernelGGL((k_ghost_fold<cplx_t<R>>), blocks_for((int64_t)nx1 * nlines, 256), 256, 0, stream(), a, nx1, nr1,
                       (int)nlines);
    CH_LAUNCH_CHECK("k_ghost_fold");
    return ST_OK;
}
template int32_t ghost_fold<double>(double2*, int, int, int64_t);
template int32_t ghost_fold<float>(float2*, int, int, int64_t);

// Mode normalisation and ghost row of the gathered field.  grid_deps.f90:219-266.
// One thread per (ix, mode, component) column, so the ghost row is written
// from the already scaled row 1 by the thread that scaled it.
__global__ void k_eb_correction(double2* __restrict__ eb, int nx1, int nr1, int nmodes, int ncomp) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx1 * nmodes * ncomp) return;
    const int64_t ix = t % nx1, q = t / nx1;
    const int m = (int)(q % nmodes);
    const double pi_inv = 1.0 / M_PI;
    const double f = m == 0 ? 0.5 * pi_inv : pi_inv;
    double2* col = eb + q * (int64_t)nx1 * nr1 + ix;
    for (int ir = 1; ir < nr1; ir++) col[(int64_t)ir * nx1] = crs(f, col[(int64_t)ir * nx1]);
    col[0] = m == 0 ? col[nx1] : cneg(col[nx1]);
}

// ----------------------------------------------------------------- devices

// devices.f90:162-203.  params = (a0, lambda, X0, Lx).
__global__ void k_undul_analytic(const double* __restrict__ coord, PStride sx, double* __restrict__ F, PStride sf,
                                 UndulParams up, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    double f[6] = {0, 0, 0, 0, 0, 0};
    undul_field(up, coord[sx.at(ip, 0)], coord[sx.at(ip, 1)], f);
    F[sf.at(ip, 4)] += f[4];
    F[sf.at(ip, 3)] += f[3];
}

// devices.f90:18-62 (TAP = 0) and 64-115 (TAP = 1): the field map a0(2, nx)
// (two rows: the amplitudes behind Fld(5) and Fld(4)) on nodes Xleft + (i-1) dx,
// i = 1..nx, interpolated with quadratic weights around node ix; particles
// outside [Xleft + dx, Xleft + (nx - 1) dx] see no field.  A node index outside
// 1..nx contributes zero (the reference reads a0(:, 0), before the map, for
// x < Xleft + 1.5 dx).
template <int TAP>
__global__ void k_undul_mapped(const double* __restrict__ coord, double* __restrict__ F,
                               const double* __restrict__ a0, int nx, double Xleft, double dx, double dx_inv,
                               double Xright, double ku, double Lx, double taper, double x_shift, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    const double xp = coord[3 * (int64_t)ip];
    if (xp < Xleft + dx || xp > Xright - dx) return;
    const double yp = coord[3 * (int64_t)ip + 1];
    const double fx = floor((xp - Xleft) * dx_inv + 0.5);
    if (!(fx >= 0.0 && fx <= (double)nx + 1.0)) return;      // NaN or off the map: no node in range
    const int ix = (int)fx;
    const double ddx = (xp - Xleft) * dx_inv - ix;
    const double s[3] = {0.5 * ((0.5 - ddx) * (0.5 - ddx)), 0.75 - ddx * ddx, 0.5 * ((0.5 + ddx) * (0.5 + ddx))};
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int j = ix - 1 + k;                               // 1-based node
        const double v1 = (j >= 1 && j <= nx) ? a0[2 * (int64_t)(j - 1)] : 0.0;
        const double v2 = (j >= 1 && j <= nx) ? a0[2 * (int64_t)(j - 1) + 1] : 0.0;
        b1 = b1 + s[k] * v1;
        b2 = b2 + s[k] * v2;
    }
    if (TAP) {
        const double amp = 1.0 + taper * (xp - x_shift - Xleft - 0.5 * Lx) / (0.5 * Lx);
        F[6 * (int64_t)ip + 4] += amp * b1 * cosh(ku * yp);
        F[6 * (int64_t)ip + 3] += amp * b2 * sinh(ku * yp);
    } else {
        F[6 * (int64_t)ip + 4] += b1 * cosh(ku * yp);
        F[6 * (int64_t)ip + 3] += b2 * sinh(ku * yp);
    }
}

template <int TAP>
static int32_t undul_mapped_impl(const double* coord, double* Fld, const double* a0, const double* params, int32_t np,
                                 int32_t nx) {
    constexpr int NP = TAP ? 5 : 3;
    if (np < 0 || nx < 1 || params == nullptr || a0 == nullptr) return fail(ST_EINVAL, "undul_mapped: bad arguments");
    if (np == 0) return ST_OK;
    double hp[NP];
    if (is_device_ptr(params)) CH_HIP(hipMemcpy(hp, params, sizeof hp, hipMemcpyDeviceToHost));
    else memcpy(hp, params, sizeof hp);
    const double lambda = hp[0], Xleft = hp[1], dx = hp[2];
    const double Lx = TAP ? hp[3] : 0.0, taper = TAP ? hp[4] : 0.0;
    Stage st;
    const double* x = st.in(coord, 3 * (size_t)np);
    const double* a = st.in(a0, 2 * (size_t)nx);
    double* F = st.inout(Fld, 6 * (size_t)np);
    CH_TRY(st.status());
    hipLaunchKernelGGL((k_undul_mapped<TAP>), blocks_for(np, 256), 256, 0, stream(), x, F, a, nx, Xleft, dx, 1.0 / dx,
                       Xleft + nx * dx, 2.0 * M_PI / lambda, Lx, taper, 0.5 * (nx * dx - Lx), np);
    CH_LAUNCH_CHECK("k_undul_mapped");
    return st.finish();
}

// devices.f90:205-252: a plane wave of amplitude a0 and wavelength lambda at
// angle theta to x, phase phi0, with linear ramps of length `ramp` at both ends of
// [X0, X0 + Lx]; adds to Fld(3), Fld(4), Fld(5) (0-based 2, 3, 4).
__global__ void k_planewave(const double* __restrict__ coord, double* __restrict__ F, double t, double a0,
                            double X0, double Lx, double ramp, double sinth, double costh, double k0, double phi0,
                            int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    const double x = coord[3 * (int64_t)ip], y = coord[3 * (int64_t)ip + 1];
    double ampl;
    if (x <= X0 || x >= X0 + Lx) ampl = 0.0;
    else if (x > X0 && x < X0 + ramp) ampl = (x - X0) / ramp;
    else if (x > X0 + Lx - ramp && x < X0 + Lx) ampl = (X0 + Lx - x) / ramp;
    else ampl = 1.0;
    ampl = a0 * ampl * sin(k0 * (x * costh + y * sinth - t) + phi0);
    double* f = F + 6 * (int64_t)ip;
    f[2] = f[2] + ampl;
    f[3] = f[3] + ampl * sinth;
    f[4] = f[4] - ampl * costh;
}

// devices.f90:254-297: a Gaussian pulse centred at (x0 + axis t, y0, z0) with
// rms-like lengths (Lx, Ly, Lz) and carrier sin(2 pi xp / lambda); adds E to
// Fld(3) and -axis E to Fld(5).
__global__ void k_gaussbeam(const double* __restrict__ coord, double* __restrict__ F, double time, double a0,
                            double lambda, double axis, double x0, double y0, double z0, double Lx2_inv,
                            double Ly2_inv, double Lz2_inv, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    const double* c = coord + 3 * (int64_t)ip;
    const double xp = c[0] - x0 - axis * time, yp = c[1] - y0, zp = c[2] - z0;
    const double e = a0 * exp(-(xp * xp * Lx2_inv) - yp * yp * Ly2_inv - zp * zp * Lz2_inv) *
                     sin(2.0 * M_PI / lambda * xp);
    double* f = F + 6 * (int64_t)ip;
    f[2] = f[2] + e;
    f[4] = f[4] - axis * e;
}

// ----------------------------------------------------------------- lattice

// particle_tools.f90:84-128.  One thread per (ix, ir, ip) lattice slot.
// Whether a slot survives the reference's rc <= 0 test depends on (ir, ip)
// only, so the host passes rank[ir*PPC+ip] (the slot's position among the
// surviving slots of its radial row, -1 if dropped), cnt[ir] (survivors per
// cell of that row) and base[ir] (first output index of the row); the output
// order is then the reference's (ir outer, ix, ip inner).
__global__ void k_genparts(double* __restrict__ coord, const double* __restrict__ Xgrid,
                           const double* __restrict__ Rgrid, const double* __restrict__ RandPackO,
                           const double* __restrict__ PackX, const double* __restrict__ PackR,
                           const double2* __restrict__ PackO, const int64_t* __restrict__ tab, int nx, int nr,
                           int PPC) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nslot = (int64_t)(nx - 1) * (nr - 1) * PPC;
    if (t >= nslot) return;
    const int ip = (int)(t % PPC);
    const int64_t cell = t / PPC;
    const int ix = (int)(cell % (nx - 1)), ir = (int)(cell / (nx - 1));
    const int64_t* rank = tab;
    const int64_t* cnt = tab + (int64_t)(nr - 1) * PPC;
    const int64_t* base = cnt + (nr - 1);
    const int64_t rk = rank[(int64_t)ir * PPC + ip];
    if (rk < 0) return;
    const int64_t o = base[ir] + (int64_t)ix * cnt[ir] + rk;
    const double dr_2 = 0.5 * (Rgrid[1] - Rgrid[0]);
    const double r0 = Rgrid[ir] + dr_2, r1 = Rgrid[ir + 1] + dr_2;
    const double x0 = Xgrid[ix], x1 = Xgrid[ix + 1];
    const double ang = 2.0 * M_PI * RandPackO[ix + (int64_t)nx * ir];
    const double2 O = cmul(PackO[ip], cmk(cos(ang), sin(ang)));
    const double xc = x0 + (x1 - x0) * PackX[ip];
    const double rc = r0 + (r1 - r0) * PackR[ip];
    coord[4 * o + 0] = xc;
    coord[4 * o + 1] = rc * O.y;
    coord[4 * o + 2] = rc * O.x;
    coord[4 * o + 3] = rc;
}

// Launch wrappers shared with the resident entry points (pic_resident.hip).
int32_t launch_push_coords(double* x, PStride sx, const double* u, PStride su, double* xc, PStride sc, double dt,
                           int np) {
    KTimer kt(KF_PUSH_COORDS);
    hipLaunchKernelGGL(k_push_coords, blocks_for(np, 256), 256, 0, stream(), x, sx, u, su, xc, sc, dt, np);
    CH_LAUNCH_CHECK("k_push_coords");
    return ST_OK;
}

int32_t launch_dep_curr(const double* x, PStride sx, const double* u, PStride su, const double* w, double2* curr,
                        const GridGeom& g, int ip0, int ip1) {
    if (ip1 <= ip0) return ST_OK;
    KTimer kt(KF_DEP_CURR);
    hipLaunchKernelGGL(k_dep_curr, blocks_for(ip1 - ip0, 256), 256, 0, stream(), x, sx, u, su, w, curr, g, ip0, ip1);
    CH_LAUNCH_CHECK("k_dep_curr");
    return ST_OK;
}

template <typename W>
int32_t launch_dep_dens(const double* x, PStride sx, const W* w, double2* dens, const GridGeom& g, int ip0, int ip1) {
    if (ip1 <= ip0) return ST_OK;
    KTimer kt(KF_DEP_DENS);
    hipLaunchKernelGGL(k_dep_dens<W>, blocks_for(ip1 - ip0, 256), 256, 0, stream(), x, sx, w, dens, g, ip0, ip1);
    CH_LAUNCH_CHECK("k_dep_dens");
    return ST_OK;
}
template int32_t launch_dep_dens<double>(const double*, PStride, const double*, double2*, const GridGeom&, int, int);
template int32_t launch_dep_dens<float>(const double*, PStride, const float*, double2*, const GridGeom&, int, int);

// Density on the grid of get_dens_on_grid (species.py:400-427) after the
// deposit and its ghost fold: each radial row ir times VGrid[ir] (zero at the
// ghost row), in place on dens(nx1, nr1, nmodes).
__global__ void k_scale_rows(double2* __restrict__ a, const double* __restrict__ v, int nx1, int nr1, int nlines) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx1 * nr1 * nlines) return;
    const int ir = (int)((t / nx1) % nr1);
    a[t] = crs(v[ir], a[t]);
}

int32_t scale_rows(double2* a, const double* v, int nx1, int nr1, int nlines) {
    hipLaunchKernelGGL(k_scale_rows, blocks_for((int64_t)nx1 * nr1 * nlines, 256), 256, 0, stream(), a, v, nx1, nr1,
                       nlines);
    CH_LAUNCH_CHECK("k_scale_rows");
    return ST_OK;
}

static int32_t check_modes(int nkO) {
    if (nkO < 0 || nkO >= MAX_MODES)
        return fail(ST_EUNSUP, "azimuthal modes 0..%d requested; this build supports up to %d modes", nkO, MAX_MODES);
    return ST_OK;
}

}  // namespace chimera

using namespace chimera;

static const PStride AOS3 = {3, 1}, AOS6 = {6, 1};

extern "C" {

int32_t chimera_push_velocs(double* momenta, const double* Fld, double dt, int32_t np) {
    if (np < 0) return fail(ST_EINVAL, "push_velocs: np < 0");
    if (np == 0) return ST_OK;
    Stage st;
    double* u = st.inout(momenta, 3 * (size_t)np);
    const double* F = st.in(Fld, 6 * (size_t)np);
    CH_TRY(st.status());
    hipLaunchKernelGGL(k_push_velocs, blocks_for(np, 256), 256, 0, stream(), u, AOS3, F, AOS6, dt, np);
    CH_LAUNCH_CHECK("k_push_velocs");
    return st.finish();
}

int32_t chimera_push_coords(double* coord, const double* momenta, double* coord_cntr, double dt, int32_t np) {
    if (np < 0) return fail(ST_EINVAL, "push_coords: np < 0");
    if (np == 0) return ST_OK;
    Stage st;
    double* x = st.inout(coord, 3 * (size_t)np);
    const double* u = st.in(momenta, 3 * (size_t)np);
    double* xc = st.inout(coord_cntr, 3 * (size_t)np);
    CH_TRY(st.status());
    hipLaunchKernelGGL(k_push_coords, blocks_for(np, 256), 256, 0, stream(), x, AOS3, u, AOS3, xc, AOS3, dt, np);
    CH_LAUNCH_CHECK("k_push_coords");
    return st.finish();
}

int32_t chimera_proj_fld(const double* coord, const double* wghts, const double* Fld, double* Fld_tot, double leftX,
                         const double* Rgrid, double dx_inv, double dr_inv, int32_t np, int32_t nx, int32_t nr,
                         int32_t nkO) {
    if (np < 0 || nx < 1 || nr < 1) return fail(ST_EINVAL, "proj_fld: bad sizes np=%d nx=%d nr=%d", np, nx, nr);
    CH_TRY(check_modes(nkO));
    if (np == 0) return ST_OK;
    const size_t ng = (size_t)(nx + 1) * (nr + 1) * (nkO + 1) * 6;
    Stage st;
    const double* x = st.in(coord, 3 * (size_t)np);
    const double* w = st.in(wghts, (size_t)np);
    const double2* F = reinterpret_cast<const double2*>(st.in(Fld, 2 * ng));
    double* Ft = st.inout(Fld_tot, 6 * (size_t)np);
    const double* R = st.in(Rgrid, (size_t)nr + 1);
    CH_TRY(st.status());
    GridGeom g = make_geom(leftX, R, dx_inv, dr_inv, nx, nr, nkO);
    CH_TRY(load_rgrid_edges(g, Rgrid, R));
    hipLaunchKernelGGL(k_proj_fld, blocks_for(np, 256), 256, 0, stream(), x, AOS3, w, F, Ft, AOS6, g, np);
    CH_LAUNCH_CHECK("k_proj_fld");
    return st.finish();
}

static int32_t dep_curr_impl(const double* coord, const double* momenta, const double* wghts, double* curr,
                             double leftX, const double* Rgrid, double dx_inv, double dr_inv, int32_t np, int32_t nx,
                             int32_t nr, int32_t nkO, int32_t ip0, int32_t ip1) {
    if (np < 0 || nx < 1 || nr < 1) return fail(ST_EINVAL, "dep_curr: bad sizes np=%d nx=%d nr=%d", np, nx, nr);
    CH_TRY(check_modes(nkO));
    const size_t ng = (size_t)(nx + 1) * (nr + 1) * (nkO + 1) * 3;
    Stage st;
    const double* x = st.in(coord, 3 * (size_t)np);
    const double* u = st.in(momenta, 3 * (size_t)np);
    const double* w = st.in(wghts, (size_t)np);
    double2* J = reinterpret_cast<double2*>(st.inout(curr, 2 * ng));
    const double* R = st.in(Rgrid, (size_t)nr + 1);
    CH_TRY(st.status());
    GridGeom g = make_geom(leftX, R, dx_inv, dr_inv, nx, nr, nkO);
    CH_TRY(load_rgrid_edges(g, Rgrid, R));
    if (ip1 > ip0) {
        hipLaunchKernelGGL(k_dep_curr, blocks_for(ip1 - ip0, 256), 256, 0, stream(), x, AOS3, u, AOS3, w, J, g, ip0,
                           ip1);
        CH_LAUNCH_CHECK("k_dep_curr");
    }
    const int nlines = (nkO + 1) * 3;
    hipLaunchKernelGGL(k_ghost_fold<double2>, blocks_for((int64_t)(nx + 1) * nlines, 256), 256, 0, stream(), J, nx + 1, nr + 1,
                       nlines);
    CH_LAUNCH_CHECK("k_ghost_fold");
    return st.finish();
}

static int32_t dep_dens_impl(const double* coord, const double* wghts, double* dens, double leftX,
                             const double* Rgrid, double dx_inv, double dr_inv, int32_t np, int32_t nx, int32_t nr,
                             int32_t nkO, int32_t ip0, int32_t ip1) {
    if (np < 0 || nx < 1 || nr < 1) return fail(ST_EINVAL, "dep_dens: bad sizes np=%d nx=%d nr=%d", np, nx, nr);
    CH_TRY(check_modes(nkO));
    const size_t ng = (size_t)(nx + 1) * (nr + 1) * (nkO + 1);
    Stage st;
    const double* x = st.in(coord, 3 * (size_t)np);
    const double* w = st.in(wghts, (size_t)np);
    double2* D = reinterpret_cast<double2*>(st.inout(dens, 2 * ng));
    const double* R = st.in(Rgrid, (size_t)nr + 1);
    CH_TRY(st.status());
    GridGeom g = make_geom(leftX, R, dx_inv, dr_inv, nx, nr, nkO);
    CH_TRY(load_rgrid_edges(g, Rgrid, R));
    if (ip1 > ip0) {
        hipLaunchKernelGGL(k_dep_dens<double>, blocks_for(ip1 - ip0, 256), 256, 0, stream(), x, AOS3, w, D, g, ip0,
                           ip1);
        CH_LAUNCH_CHECK("k_dep_dens");
    }
    hipLaunchKernelGGL(k_ghost_fold<double2>, blocks_for((int64_t)(nx + 1) * (nkO + 1), 256), 256, 0, stream(), D, nx + 1,
                       nr + 1, nkO + 1);
    CH_LAUNCH_CHECK("k_ghost_fold");
    return st.finish();
}

int32_t chimera_dep_curr(const double* coord, const double* momenta, const double* wghts, double* curr, double leftX,
                         const double* Rgrid, double dx_inv, double dr_inv, int32_t np, int32_t nx, int32_t nr,
                         int32_t nkO) {
    return dep_curr_impl(coord, momenta, wghts, curr, leftX, Rgrid, dx_inv, dr_inv, np, nx, nr, nkO, 0, np);
}

int32_t chimera_dep_curr_chnk(const double* coord, const double* momenta, const double* wghts, double* curr,
                              const int32_t* IndInChunk, int32_t guards, double leftX, const double* Rgrid,
                              double dx_inv, double dr_inv, int32_t np, int32_t nx, int32_t nr, int32_t nkO,
                              int32_t nchnk) {
    (void)guards;
    if (nchnk < 1 || IndInChunk == nullptr) return fail(ST_EINVAL, "dep_curr_chnk: nchnk < 1");
    int32_t lo = 0, hi = 0;
    if (is_device_ptr(IndInChunk)) {
        CH_HIP(hipMemcpy(&lo, IndInChunk, sizeof lo, hipMemcpyDeviceToHost));
        CH_HIP(hipMemcpy(&hi, IndInChunk + nchnk, sizeof hi, hipMemcpyDeviceToHost));
    } else {
        lo = IndInChunk[0];
        hi = IndInChunk[nchnk];
    }
    if (lo < 0 || hi > np || lo > hi) return fail(ST_EINVAL, "dep_curr_chnk: chunk bounds [%d,%d) outside np=%d", lo, hi, np);
    return dep_curr_impl(coord, momenta, wghts, curr, leftX, Rgrid, dx_inv, dr_inv, np, nx, nr, nkO, lo, hi);
}

int32_t chimera_dep_dens(const double* coord, const double* wghts, double* dens, double leftX, const double* Rgrid,
                         double dx_inv, double dr_inv, int32_t np, int32_t nx, int32_t nr, int32_t nkO) {
    return dep_dens_impl(coord, wghts, dens, leftX, Rgrid, dx_inv, dr_inv, np, nx, nr, nkO, 0, np);
}

int32_t chimera_dep_dens_chnk(const double* coord, const double* wghts, double* dens, const int32_t* IndInChunk,
                              int32_t guards, double leftX, const double* Rgrid, double dx_inv, double dr_inv,
                              int32_t np, int32_t nx, int32_t nr, int32_t nkO, int32_t nchnk) {
    (void)guards;
    if (nchnk < 1 || IndInChunk == nullptr) return fail(ST_EINVAL, "dep_dens_chnk: nchnk < 1");
    int32_t lo = 0, hi = 0;
    if (is_device_ptr(IndInChunk)) {
        CH_HIP(hipMemcpy(&lo, IndInChunk, sizeof lo, hipMemcpyDeviceToHost));
        CH_HIP(hipMemcpy(&hi, IndInChunk + nchnk, sizeof hi, hipMemcpyDeviceToHost));
    } else {
        lo = IndInChunk[0];
        hi = IndInChunk[nchnk];
    }
    if (lo < 0 || hi > np || lo > hi) return fail(ST_EINVAL, "dep_dens_chnk: chunk bounds [%d,%d) outside np=%d", lo, hi, np);
    return dep_dens_impl(coord, wghts, dens, leftX, Rgrid, dx_inv, dr_inv, np, nx, nr, nkO, lo, hi);
}

int32_t chimera_eb_correction(double* eb_spc, int32_t nx, int32_t nr, int32_t nkO) {
    if (nx < 0 || nr < 1 || nkO < 1) return fail(ST_EINVAL, "eb_correction: bad sizes nx=%d nr=%d nkO=%d", nx, nr, nkO);
    const size_t ng = (size_t)(nx + 1) * (nr + 1) * nkO * 6;
    Stage st;
    double2* E = reinterpret_cast<double2*>(st.inout(eb_spc, 2 * ng));
    CH_TRY(st.status());
    hipLaunchKernelGGL(k_eb_correction, blocks_for((int64_t)(nx + 1) * nkO * 6, 256), 256, 0, stream(), E, nx + 1,
                       nr + 1, nkO, 6);
    CH_LAUNCH_CHECK("k_eb_correction");
    return st.finish();
}

int32_t chimera_undul_analytic(const double* coord, double* Fld, double t, const double* params, int32_t np) {
    (void)t;
    if (np < 0 || params == nullptr) return fail(ST_EINVAL, "undul_analytic: bad arguments");
    if (np == 0) return ST_OK;
    double hp[4];
    if (is_device_ptr(params)) CH_HIP(hipMemcpy(hp, params, sizeof hp, hipMemcpyDeviceToHost));
    else memcpy(hp, params, sizeof hp);
    Stage st;
    const double* x = st.in(coord, 3 * (size_t)np);
    double* F = st.inout(Fld, 6 * (size_t)np);
    CH_TRY(st.status());
    const UndulParams up = make_undul(hp);
    hipLaunchKernelGGL(k_undul_analytic, blocks_for(np, 256), 256, 0, stream(), x, AOS3, F, AOS6, up, np);
    CH_LAUNCH_CHECK("k_undul_analytic");
    return st.finish();
}

int32_t chimera_undul_analytic_taper(const double* coord, double* Fld, double t, const double* params, int32_t np) {
    (void)t;
    if (np < 0 || params == nullptr) return fail(ST_EINVAL, "undul_analytic_taper: bad arguments");
    if (np == 0) return ST_OK;
    double hp[5];
    if (is_device_ptr(params)) CH_HIP(hipMemcpy(hp, params, sizeof hp, hipMemcpyDeviceToHost));
    else memcpy(hp, params, sizeof hp);
    Stage st;
    const double* x = st.in(coord, 3 * (size_t)np);
    double* F = st.inout(Fld, 6 * (size_t)np);
    CH_TRY(st.status());
    const UndulParams up = make_undul_taper(hp);
    hipLaunchKernelGGL(k_undul_analytic, blocks_for(np, 256), 256, 0, stream(), x, AOS3, F, AOS6, up, np);
    CH_LAUNCH_CHECK("k_undul_analytic");
    return st.finish();
}

int32_t chimera_undul_mapped(const double* coord, double* Fld, double t, const double* a0, const double* params,
                             int32_t np, int32_t nx) {
    (void)t;
    return undul_mapped_impl<0>(coord, Fld, a0, params, np, nx);
}

int32_t chimera_undul_mapped_tap(const double* coord, double* Fld, double t, const double* a0, const double* params,
                                 int32_t np, int32_t nx) {
    (void)t;
    return undul_mapped_impl<1>(coord, Fld, a0, params, np, nx);
}

int32_t chimera_planewave(const double* coord, double* Fld, double t, const double* params, int32_t np) {
    if (np < 0 || params == nullptr) return fail(ST_EINVAL, "planewave: bad arguments");
    if (np == 0) return ST_OK;
    double p[7];
    if (is_device_ptr(params)) CH_HIP(hipMemcpy(p, params, sizeof p, hipMemcpyDeviceToHost));
    else memcpy(p, params, sizeof p);
    Stage st;
    const double* x = st.in(coord, 3 * (size_t)np);
    double* F = st.inout(Fld, 6 * (size_t)np);
    CH_TRY(st.status());
    hipLaunchKernelGGL(k_planewave, blocks_for(np, 256), 256, 0, stream(), x, F, t, p[0], p[2], p[3], p[4], sin(p[5]),
                       cos(p[5]), 2.0 * M_PI / p[1], p[6], np);
    CH_LAUNCH_CHECK("k_planewave");
    return st.finish();
}

int32_t chimera_gaussbeam(const double* coord, double* Fld, double time, double a0, const double* params, int32_t np) {
    if (np < 0 || params == nullptr) return fail(ST_EINVAL, "gaussbeam: bad arguments");
    if (np == 0) return ST_OK;
    double p[8];
    if (is_device_ptr(params)) CH_HIP(hipMemcpy(p, params, sizeof p, hipMemcpyDeviceToHost));
    else memcpy(p, params, sizeof p);
    Stage st;
    const double* x = st.in(coord, 3 * (size_t)np);
    double* F = st.inout(Fld, 6 * (size_t)np);
    CH_TRY(st.status());
    hipLaunchKernelGGL(k_gaussbeam, blocks_for(np, 256), 256, 0, stream(), x, F, time, a0, p[0], p[1], p[2], p[3], p[4],
                       1.0 / (p[5] * p[5]), 1.0 / (p[6] * p[6]), 1.0 / (p[7] * p[7]), np);
    CH_LAUNCH_CHECK("k_gaussbeam");
    return st.finish();
}

int32_t chimera_genparts(double* coord, int32_t* indPart, const double* Xgrid, const double* Rgrid,
                         const double* RandPackO, const double* PackX, const double* PackR, const double* PackO,
                         int32_t np, int32_t nx, int32_t nr, int32_t PPC) {
    if (np < 0 || nx < 1 || nr < 1 || PPC < 1 || indPart == nullptr) return fail(ST_EINVAL, "genparts: bad arguments");
    // Host side: which slots survive the rc <= 0 test, and where they go.
    std::vector<double> hX(nx), hR(nr), hPR(PPC);
    auto fetch = [](double* dst, const double* src, size_t n) -> int32_t {
        if (is_device_ptr(src)) CH_HIP(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost));
        else memcpy(dst, src, n * sizeof(double));
        return ST_OK;
    };
    CH_TRY(fetch(hX.data(), Xgrid, nx));
    CH_TRY(fetch(hR.data(), Rgrid, nr));
    CH_TRY(fetch(hPR.data(), PackR, PPC));
    const int64_t nslot = (nx > 1 && nr > 1) ? (int64_t)(nx - 1) * (nr - 1) * PPC : 0;
    const int nrow = nr > 1 ? nr - 1 : 0;
    std::vector<int64_t> tab((size_t)nrow * PPC + 2 * (size_t)nrow + 1, 0);
    int64_t* rank = tab.data();
    int64_t* cnt = rank + (size_t)nrow * PPC;
    int64_t* base = cnt + nrow;
    const double dr_2 = nr > 1 ? 0.5 * (hR[1] - hR[0]) : 0.0;
    int64_t n = 0;
    for (int ir = 0; ir < nrow; ir++) {
        int64_t c = 0;
        for (int ip = 0; ip < PPC; ip++) {
            const double r0 = hR[ir] + dr_2, r1 = hR[ir + 1] + dr_2;
            const double rc = r0 + (r1 - r0) * hPR[ip];
            rank[(size_t)ir * PPC + ip] = (rc <= 0) ? -1 : c++;
        }
        cnt[ir] = c;
        base[ir] = n;
        n += c * (nx - 1);
    }
    if (n > np) return fail(ST_EINVAL, "genparts: %lld particles do not fit coord(4,%d)", (long long)n, np);
    Stage st;
    double* C = st.out(coord, 4 * (size_t)np);
    const double* X = st.in(Xgrid, nx);
    const double* R = st.in(Rgrid, nr);
    const double* RP = st.in(RandPackO, (size_t)nx * nr);
    const double* PX = st.in(PackX, PPC);
    const double* PRd = st.in(PackR, PPC);
    const double2* PO = reinterpret_cast<const double2*>(st.in(PackO, 2 * (size_t)PPC));
    CH_TRY(st.status());
    const int64_t* dtab = static_cast<const int64_t*>(device_table(tab.data(), tab.size() * sizeof(int64_t)));
    if (!dtab) return ST_ENOMEM;
    CH_HIP(hipMemsetAsync(C, 0, 4 * (size_t)np * sizeof(double), stream()));
    if (nslot > 0) {
        hipLaunchKernelGGL(k_genparts, blocks_for(nslot, 256), 256, 0, stream(), C, X, R, RP, PX, PRd, PO, dtab, nx,
                           nr, PPC);
        CH_LAUNCH_CHECK("k_genparts");
    }
    *indPart = (int32_t)n;
    return st.finish();
}

}  // extern "C"
