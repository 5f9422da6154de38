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

// f2py layouts of the drop-in entry points: coord (3, np), Fld (6, np)
static const PStride AOS3 = {3, 1}, AOS6 = {6, 1};

// Every device entry point (devices.f90): one thread per particle adds the
// device's field (the shared ext_device_field) to the particle's row of Fld.
__global__ void k_ext_device(const double* __restrict__ coord, PStride sx, double* __restrict__ F, PStride sf,
                             ExtDevice d, double t, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    ExtDevices D;
    D.d[0] = d;
    D.n = 1;
    D.t = t;
    double f[6];
    for (int c = 0; c < 6; c++) f[c] = F[sf.at(ip, c)];
    ext_fields<double>(D, coord[sx.at(ip, 0)], coord[sx.at(ip, 1)], coord[sx.at(ip, 2)], f);
    for (int c = 2; c < 5; c++) F[sf.at(ip, c)] = f[c];
}

int32_t make_ext_device(ExtDevice& d, int kind, const double* p, int np, const double* map, int nmap, double a0) {
    memset(&d, 0, sizeof d);
    d.kind = kind;
    static const int need[6] = {4, 5, 3, 5, 7, 8};
    if (kind < 0 || kind > DEV_GAUSSBEAM) return fail(ST_EINVAL, "external device kind %d", kind);
    if (p == nullptr || np < need[kind])
        return fail(ST_EINVAL, "external device kind %d: %d parameters, %d needed", kind, np, need[kind]);
    switch (kind) {
        case DEV_UNDUL:
        case DEV_UNDUL_TAPER:
            d.p[0] = p[0];
            d.p[1] = p[1];
            d.p[2] = p[2];
            d.p[3] = p[3];
            d.p[4] = 2.0 * M_PI / p[1];
            d.p[5] = kind == DEV_UNDUL_TAPER ? p[4] : 0.0;
            break;
        case DEV_MAPPED:
        case DEV_MAPPED_TAP: {
            if (map == nullptr || nmap < 1) return fail(ST_EINVAL, "mapped undulator: no map");
            const double lambda = p[0], Xleft = p[1], dx = p[2];
            const double Lx = kind == DEV_MAPPED_TAP ? p[3] : 0.0, taper = kind == DEV_MAPPED_TAP ? p[4] : 0.0;
            d.p[0] = Xleft;
            d.p[1] = dx;
            d.p[2] = 1.0 / dx;
            d.p[3] = Xleft + nmap * dx;
            d.p[4] = 2.0 * M_PI / lambda;
            d.p[5] = Lx;
            d.p[6] = taper;
            d.p[7] = 0.5 * (nmap * dx - Lx);
            d.map = map;
            d.nmap = nmap;
            break;
        }
        case DEV_PLANEWAVE:
            // params = (a0, lambda, X0, Lx, ramp, theta, phi0)
            d.p[0] = p[0];
            d.p[1] = p[2];
            d.p[2] = p[3];
            d.p[3] = p[4];
            d.p[4] = sin(p[5]);
            d.p[5] = cos(p[5]);
            d.p[6] = 2.0 * M_PI / p[1];
            d.p[7] = p[6];
            break;
        default:
            // params = (lambda, axis, x0, y0, z0, Lx, Ly, Lz); amplitude a0
            d.p[0] = a0;
            for (int k = 0; k < 5; k++) d.p[1 + k] = p[k];
            d.p[6] = 1.0 / (p[5] * p[5]);
            d.p[7] = 1.0 / (p[6] * p[6]);
            d.p[8] = 1.0 / (p[7] * p[7]);
            break;
    }
    return ST_OK;
}

// A drop-in device call: Fld(6, np) += the device's field at coord(3, np).
static int32_t ext_device_call(int kind, const double* coord, double* Fld, double t, const double* params, int npar,
                               const double* a0map, int nmap, double a0, int32_t np) {
    if (np < 0 || params == nullptr) return fail(ST_EINVAL, "device kind %d: bad arguments", kind);
    if (np == 0) return ST_OK;
    double hp[8];
    if (is_device_ptr(params)) CH_HIP(hipMemcpy(hp, params, npar * sizeof(double), hipMemcpyDeviceToHost));
    else memcpy(hp, params, npar * sizeof(double));
    Stage st;
    const double* x = st.in(coord, 3 * (size_t)np);
    const double* m = a0map ? st.in(a0map, 2 * (size_t)nmap) : nullptr;
    double* F = st.inout(Fld, 6 * (size_t)np);
    CH_TRY(st.status());
    ExtDevice d;
    CH_TRY(make_ext_device(d, kind, hp, npar, m, nmap, a0));
    hipLaunchKernelGGL(k_ext_device, blocks_for(np, 256), 256, 0, stream(), x, AOS3, F, AOS6, d, t, np);
    CH_LAUNCH_CHECK("k_ext_device");
    return st.finish();
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
    return ext_device_call(DEV_UNDUL, coord, Fld, t, params, 4, nullptr, 0, 0.0, np);
}

int32_t chimera_undul_analytic_taper(const double* coord, double* Fld, double t, const double* params, int32_t np) {
    return ext_device_call(DEV_UNDUL_TAPER, coord, Fld, t, params, 5, nullptr, 0, 0.0, np);
}

int32_t chimera_undul_mapped(const double* coord, double* Fld, double t, const double* a0, const double* params,
                             int32_t np, int32_t nx) {
    if (nx < 1 || a0 == nullptr) return fail(ST_EINVAL, "undul_mapped: bad arguments");
    return ext_device_call(DEV_MAPPED, coord, Fld, t, params, 3, a0, nx, 0.0, np);
}

int32_t chimera_undul_mapped_tap(const double* coord, double* Fld, double t, const double* a0, const double* params,
                                 int32_t np, int32_t nx) {
    if (nx < 1 || a0 == nullptr) return fail(ST_EINVAL, "undul_mapped_tap: bad arguments");
    return ext_device_call(DEV_MAPPED_TAP, coord, Fld, t, params, 5, a0, nx, 0.0, np);
}

int32_t chimera_planewave(const double* coord, double* Fld, double t, const double* params, int32_t np) {
    return ext_device_call(DEV_PLANEWAVE, coord, Fld, t, params, 7, nullptr, 0, 0.0, np);
}

int32_t chimera_gaussbeam(const double* coord, double* Fld, double time, double a0, const double* params, int32_t np) {
    return ext_device_call(DEV_GAUSSBEAM, coord, Fld, time, params, 8, nullptr, 0, a0, np);
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
