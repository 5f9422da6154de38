// envelope.hip -- the particle side of the envelope (KxShift) solver: the
// deposits and gather of grid_deps_env.f90 (and the chunked deposits of
// grid_deps_env_chnk.f90, with the unchunked semantics) and eb_correction_env.
//
// Envelope grids carry modes -nkO..nkO (mode m at index m + nkO).  A
// particle's weight carries the carrier e^{-i kx0 x} into the deposits and
// e^{+i kx0 x} into the gather; mode -m takes the conjugate azimuthal phase.
// What the reference does and these kernels mirror:
//  * dep_curr_env deposits only the third current component (its `do l=3,3`,
//    grid_deps_env.f90:72), with u/gamma and the complex weight;
//  * dep_dens_env multiplies by the complex weight twice
//    (grid_deps_env.f90:145-147);
//  * the deposit phase is 0 on the axis, the gather phase 1
//    (grid_deps_env.f90:56-60, :203-207);
//  * eb_correction_env scales every mode by 1/pi and sets the ghost row to
//    +row 1 for nkO = 0 and to -row 1 otherwise.
// Sums use hardware FP64 atomics (drop-in tier: every array may be host or
// device memory).  Particles whose cell is outside the arrays are skipped, as
// in the plain kernels.
#include "common.hpp"
#include "pic_device.hpp"
#include "../../include/chimera_hip.h"

namespace chimera {

// Cell, weights and phase powers ((y - i z)/r for the deposits, (y + i z)/r
// for the gather) of an envelope particle.
__device__ __forceinline__ bool env_geom(PartGeom& pg, double xp, double yp, double zp, const GridGeom& g,
                                         bool gather) {
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
    double2 p1;
    if (rp > 0.0) p1 = gather ? cmk(yp / rp, zp / rp) : cmk(yp / rp, -zp / rp);
    else p1 = gather ? cmk(1.0, 0.0) : cmk(0.0, 0.0);
    pg.ph[0] = cmk(1.0, 0.0);
#pragma unroll
    for (int m = 1; m < MAX_MODES; m++)
        if (m <= g.nkO) pg.ph[m] = cmul(pg.ph[m - 1], p1);
    return true;
}

__device__ __forceinline__ void atomic_add_c(double2* p, double2 v) {
    double* d = reinterpret_cast<double*>(p);
    unsafeAtomicAdd(d, v.x);
    unsafeAtomicAdd(d + 1, v.y);
}

// grid_deps_env.f90:34-87 (particle loop of dep_curr_env)
__global__ void k_dep_curr_env(const double* __restrict__ coord, PStride sx, const double* __restrict__ u, PStride su,
                               const double* __restrict__ w, double2* __restrict__ curr, GridGeom g, double kx0,
                               int ip0, int ip1) {
    const int ip = ip0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= ip1) return;
    g = frame_geom(g);
    const double wr = w[ip];
    if (wr == 0.0) return;
    double v[3] = {u[su.at(ip, 0)], u[su.at(ip, 1)], u[su.at(ip, 2)]};
    if ((fabs(v[0]) + fabs(v[1]) + fabs(v[2])) == 0.0) return;
    const double xp = coord[sx.at(ip, 0)];
    PartGeom pg;
    if (!env_geom(pg, xp, coord[sx.at(ip, 1)], coord[sx.at(ip, 2)], g, false)) return;
    const double gp = sqrt(1.0 + v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double vz = v[2] / gp;
    const double2 wp = cmul(cmk(wr, 0.0), cmk(cos(xp * kx0), -sin(xp * kx0)));
    const int nm = 2 * g.nkO + 1;
    const int64_t sr = g.nx + 1, sm = sr * (g.nr + 1), sl = sm * nm;
    double2* base = curr + pg.ix + pg.ir * sr + 2 * sl;     // component 3 only
    for (int iO = 0; iO <= g.nkO; iO++) {
        const double2 pv = cmul(pg.ph[iO], cmk(vz, 0.0));
        const double2 pc = cmul(cconj(pg.ph[iO]), cmk(vz, 0.0));
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) {
                const double2 cp = cmul(cmk(pg.S[0][a] * pg.S[1][b], 0.0), wp);
                atomic_add_c(base + a + b * sr + (g.nkO + iO) * sm, cmul(pv, cp));
                if (iO > 0) atomic_add_c(base + a + b * sr + (g.nkO - iO) * sm, cmul(pc, cp));
            }
    }
}

// grid_deps_env.f90:112-154 (particle loop of dep_dens_env); weight applied twice
template <typename W>
__global__ void k_dep_dens_env(const double* __restrict__ coord, PStride sx, const W* __restrict__ w,
                               double2* __restrict__ dens, GridGeom g, double kx0, int ip0, int ip1) {
    const int ip = ip0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= ip1) return;
    g = frame_geom(g);
    const double wr = (double)w[ip];
    if (wr == 0.0) return;
    const double xp = coord[sx.at(ip, 0)];
    PartGeom pg;
    if (!env_geom(pg, xp, coord[sx.at(ip, 1)], coord[sx.at(ip, 2)], g, false)) return;
    const double2 wp = cmul(cmk(wr, 0.0), cmk(cos(xp * kx0), -sin(xp * kx0)));
    const int64_t sr = g.nx + 1, sm = sr * (g.nr + 1);
    double2* base = dens + pg.ix + pg.ir * sr;
    for (int iO = 0; iO <= g.nkO; iO++)
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) {
                const double2 dp = cmul(cmul(cmk(pg.S[0][a] * pg.S[1][b], 0.0), wp), wp);
                atomic_add_c(base + a + b * sr + (g.nkO + iO) * sm, cmul(pg.ph[iO], dp));
                if (iO > 0) atomic_add_c(base + a + b * sr + (g.nkO - iO) * sm, cmul(cconj(pg.ph[iO]), dp));
            }
}

// grid_deps_env.f90:180-236 (proj_fld_env): Ftot(6, np) += Re(sum over modes,
// +m then -m, of S_a S_b e^{i kx0 x} phase Fld), in the reference's order
template <typename R>
__global__ void k_proj_fld_env(const double* __restrict__ coord, PStride sx, const R* __restrict__ w,
                               const double2* __restrict__ Fld, R* __restrict__ Ftot, PStride sf, GridGeom g,
                               double kx0, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    if ((double)w[ip] == 0.0) return;
    const double xp = coord[sx.at(ip, 0)];
    PartGeom pg;
    if (!env_geom(pg, xp, coord[sx.at(ip, 1)], coord[sx.at(ip, 2)], g, true)) return;
    const double2 carrier = cmk(cos(xp * kx0), sin(xp * kx0));
    const int nm = 2 * g.nkO + 1;
    const int64_t sr = g.nx + 1, sm = sr * (g.nr + 1), sl = sm * nm;
    double2 pc[2][2];
    for (int b = 0; b < 2; b++)
        for (int a = 0; a < 2; a++) pc[a][b] = cmul(cmk(pg.S[1][b] * pg.S[0][a], 0.0), carrier);
    double Fp[6] = {0, 0, 0, 0, 0, 0};
    const double2* base = Fld + pg.ix + pg.ir * sr;
    for (int iO = 0; iO <= g.nkO; iO++)
        for (int sgn = 0; sgn < (iO > 0 ? 2 : 1); sgn++) {
            const double2 ph = sgn ? cconj(pg.ph[iO]) : pg.ph[iO];
            const int m = sgn ? g.nkO - iO : g.nkO + iO;
            double2 po[2][2];
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) po[a][b] = cmul(pc[a][b], ph);
            for (int l = 0; l < 6; l++) {
                double s = 0.0;
                for (int b = 0; b < 2; b++)
                    for (int a = 0; a < 2; a++) {
                        const double2 f = base[a + b * sr + m * sm + l * sl];
                        s += po[a][b].x * f.x - po[a][b].y * f.y;
                    }
                Fp[l] = Fp[l] + s;
            }
        }
    for (int l = 0; l < 6; l++) Ftot[sf.at(ip, l)] = (R)((double)Ftot[sf.at(ip, l)] + Fp[l]);
}

// grid_deps_env.f90:240-283, one thread per (ix, mode, component) column
template <typename C>
__global__ void k_eb_correction_env(C* __restrict__ eb, int nx1, int nr1, int nm, int nkO) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx1 * nm * 6) return;
    const int64_t ix = t % nx1, q = t / nx1;
    const double pi_inv = 1.0 / M_PI;
    C* col = eb + q * (int64_t)nx1 * nr1 + ix;
    for (int ir = 0; ir < nr1; ir++) {
        C& e = col[(int64_t)ir * nx1];
        e.x = (decltype(e.x))(pi_inv * (double)e.x);
        e.y = (decltype(e.y))(pi_inv * (double)e.y);
    }
    col[0] = nkO == 0 ? col[nx1] : cneg(col[nx1]);
}

static const PStride ENV_AOS3 = {3, 1}, ENV_AOS6 = {6, 1};

static int32_t env_check(int32_t np, int32_t nx, int32_t nr, int32_t nkO, const char* what) {
    if (np < 0 || nx < 1 || nr < 1) return fail(ST_EINVAL, "%s: bad sizes np=%d nx=%d nr=%d", what, np, nx, nr);
    if (nkO < 0 || nkO >= MAX_MODES)
        return fail(ST_EUNSUP, "%s: modes -%d..%d; this build supports |m| < %d", what, nkO, nkO, MAX_MODES);
    return ST_OK;
}

// ghost fold of every mode and component: row 1 -= row 0, row 0 = 0
__global__ void k_env_fold(double2* __restrict__ a, int nx1, int nr1, int64_t nlines) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx1 * nlines) return;
    const int64_t ix = t % nx1, q = t / nx1;
    double2* r0 = a + ix + q * (int64_t)nx1 * nr1;
    r0[nx1] = csub(r0[nx1], r0[0]);
    r0[0] = cmk(0.0, 0.0);
}

static int32_t dep_curr_env_impl(const double* coord, const double* momenta, const double* wghts, double* curr,
                                 double leftX, const double* Rgrid, double dx_inv, double dr_inv, double kx0,
                                 int32_t np, int32_t nx, int32_t nr, int32_t nkO, int32_t ip0, int32_t ip1) {
    CH_TRY(env_check(np, nx, nr, nkO, "dep_curr_env"));
    const int nm = 2 * nkO + 1;
    const size_t ng = (size_t)(nx + 1) * (nr + 1) * nm * 3;
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
        hipLaunchKernelGGL(k_dep_curr_env, blocks_for(ip1 - ip0, 256), 256, 0, stream(), x, ENV_AOS3, u, ENV_AOS3, w,
                           J, g, kx0, ip0, ip1);
        CH_LAUNCH_CHECK("k_dep_curr_env");
    }
    const int64_t nlines = (int64_t)nm * 3;
    hipLaunchKernelGGL(k_env_fold, blocks_for((int64_t)(nx + 1) * nlines, 256), 256, 0, stream(), J, nx + 1, nr + 1,
                       nlines);
    CH_LAUNCH_CHECK("k_env_fold");
    return st.finish();
}

static int32_t dep_dens_env_impl(const double* coord, const double* wghts, double* dens, double leftX,
                                 const double* Rgrid, double dx_inv, double dr_inv, double kx0, int32_t np, int32_t nx,
                                 int32_t nr, int32_t nkO, int32_t ip0, int32_t ip1) {
    CH_TRY(env_check(np, nx, nr, nkO, "dep_dens_env"));
    const int nm = 2 * nkO + 1;
    const size_t ng = (size_t)(nx + 1) * (nr + 1) * nm;
    Stage st;
    const double* x = st.in(coord, 3 * (size_t)np);
    const double* w = st.in(wghts, (size_t)np);
    double2* D = reinterpret_cast<double2*>(st.inout(dens, 2 * ng));
    const double* R = st.in(Rgrid, (size_t)nr + 1);
    CH_TRY(st.status());
    GridGeom g = make_geom(leftX, R, dx_inv, dr_inv, nx, nr, nkO);
    CH_TRY(load_rgrid_edges(g, Rgrid, R));
    if (ip1 > ip0) {
        hipLaunchKernelGGL(k_dep_dens_env<double>, blocks_for(ip1 - ip0, 256), 256, 0, stream(), x, ENV_AOS3, w, D,
                           g, kx0, ip0, ip1);
        CH_LAUNCH_CHECK("k_dep_dens_env");
    }
    hipLaunchKernelGGL(k_env_fold, blocks_for((int64_t)(nx + 1) * nm, 256), 256, 0, stream(), D, nx + 1, nr + 1,
                       (int64_t)nm);
    CH_LAUNCH_CHECK("k_env_fold");
    return st.finish();
}

// ---------------------------------------------------------------- resident tier
//
// The same particle loops on structure-of-arrays particles in device memory
// (component c of particle i at c*ld + i), asynchronous on the library
// stream, deposits WITHOUT the ghost fold (the forward transform applies it,
// as in the plain resident path), and the gather fused with the external
// devices and the Boris push (chimera_main.py:130-151, :91-92): the field on
// the particle starts at zero, gets the envelope gather, the devices' fields
// in the species' order, then pushes the momenta.

template <typename R>
__global__ void k_gather_push_env(const double* __restrict__ coord, PStride sx, R* __restrict__ u, PStride su,
                                  const R* __restrict__ w, const double2* __restrict__ Fld, GridGeom g, double kx0,
                                  double push_dt, ExtDevices devs, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    g = frame_geom(g);
    const double xp = coord[sx.at(ip, 0)], yp = coord[sx.at(ip, 1)], zp = coord[sx.at(ip, 2)];
    double Fp[6] = {0, 0, 0, 0, 0, 0};
    PartGeom pg;
    if ((double)w[ip] != 0.0 && env_geom(pg, xp, yp, zp, g, true)) {
        // grid_deps_env.f90:180-236, in the order of k_proj_fld_env
        const double2 carrier = cmk(cos(xp * kx0), sin(xp * kx0));
        const int nm = 2 * g.nkO + 1;
        const int64_t sr = g.nx + 1, sm = sr * (g.nr + 1), sl = sm * nm;
        double2 pc[2][2];
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) pc[a][b] = cmul(cmk(pg.S[1][b] * pg.S[0][a], 0.0), carrier);
        const double2* base = Fld + pg.ix + pg.ir * sr;
        for (int iO = 0; iO <= g.nkO; iO++)
            for (int sgn = 0; sgn < (iO > 0 ? 2 : 1); sgn++) {
                const double2 ph = sgn ? cconj(pg.ph[iO]) : pg.ph[iO];
                const int m = sgn ? g.nkO - iO : g.nkO + iO;
                double2 po[2][2];
                for (int b = 0; b < 2; b++)
                    for (int a = 0; a < 2; a++) po[a][b] = cmul(pc[a][b], ph);
                for (int l = 0; l < 6; l++) {
                    double s = 0.0;
                    for (int b = 0; b < 2; b++)
                        for (int a = 0; a < 2; a++) {
                            const double2 f = base[a + b * sr + m * sm + l * sl];
                            s += po[a][b].x * f.x - po[a][b].y * f.y;
                        }
                    Fp[l] = Fp[l] + s;
                }
            }
    }
    R f[6];
    for (int l = 0; l < 6; l++) f[l] = (R)Fp[l];
    if (devs.n) ext_fields<R>(devs, xp, yp, zp, f);
    R m3[3] = {u[su.at(ip, 0)], u[su.at(ip, 1)], u[su.at(ip, 2)]};
    boris<R>(m3, f, (R)(0.5 * push_dt));
    for (int c = 0; c < 3; c++) u[su.at(ip, c)] = m3[c];
}

// the particle range [lo, hi) of the chunks (IndInChunk(0), IndInChunk(nchnk))
static int32_t chunk_range(const int32_t* IndInChunk, int32_t nchnk, int32_t np, int32_t* lo, int32_t* hi) {
    if (nchnk < 1 || IndInChunk == nullptr) return fail(ST_EINVAL, "chunked deposit: nchnk < 1");
    if (is_device_ptr(IndInChunk)) {
        CH_HIP(hipMemcpy(lo, IndInChunk, sizeof *lo, hipMemcpyDeviceToHost));
        CH_HIP(hipMemcpy(hi, IndInChunk + nchnk, sizeof *hi, hipMemcpyDeviceToHost));
    } else {
        *lo = IndInChunk[0];
        *hi = IndInChunk[nchnk];
    }
    if (*lo < 0 || *hi > np || *lo > *hi) return fail(ST_EINVAL, "chunked deposit: chunk range %d..%d of %d", *lo, *hi, np);
    return ST_OK;
}

}  // namespace chimera

using namespace chimera;

extern "C" {

int32_t chimera_dep_curr_env(const double* coord, const double* momenta, const double* wghts, double* curr,
                             double leftX, const double* Rgrid, double dx_inv, double dr_inv, double kx0, int32_t np,
                             int32_t nx, int32_t nr, int32_t nkO) {
    return dep_curr_env_impl(coord, momenta, wghts, curr, leftX, Rgrid, dx_inv, dr_inv, kx0, np, nx, nr, nkO, 0, np);
}

int32_t chimera_dep_dens_env(const double* coord, const double* wghts, double* dens, double leftX, const double* Rgrid,
                             double dx_inv, double dr_inv, double kx0, int32_t np, int32_t nx, int32_t nr,
                             int32_t nkO) {
    return dep_dens_env_impl(coord, wghts, dens, leftX, Rgrid, dx_inv, dr_inv, kx0, np, nx, nr, nkO, 0, np);
}

int32_t chimera_dep_curr_env_chnk(const double* coord, const double* momenta, const double* wghts, double* curr,
                                  const int32_t* IndInChunk, int32_t guards, double leftX, const double* Rgrid,
                                  double dx_inv, double dr_inv, double kx0, int32_t np, int32_t nx, int32_t nr,
                                  int32_t nkO, int32_t nchnk) {
    (void)guards;
    int32_t lo = 0, hi = 0;
    CH_TRY(chunk_range(IndInChunk, nchnk, np, &lo, &hi));
    return dep_curr_env_impl(coord, momenta, wghts, curr, leftX, Rgrid, dx_inv, dr_inv, kx0, np, nx, nr, nkO, lo, hi);
}

int32_t chimera_dep_dens_env_chnk(const double* coord, const double* wghts, double* dens, const int32_t* IndInChunk,
                                  int32_t guards, double leftX, const double* Rgrid, double dx_inv, double dr_inv,
                                  double kx0, int32_t np, int32_t nx, int32_t nr, int32_t nkO, int32_t nchnk) {
    (void)guards;
    int32_t lo = 0, hi = 0;
    CH_TRY(chunk_range(IndInChunk, nchnk, np, &lo, &hi));
    return dep_dens_env_impl(coord, wghts, dens, leftX, Rgrid, dx_inv, dr_inv, kx0, np, nx, nr, nkO, lo, hi);
}

int32_t chimera_proj_fld_env(const double* coord, const double* wghts, const double* Fld, double* Fld_tot,
                             double leftX, const double* Rgrid, double dx_inv, double dr_inv, double kx0, int32_t np,
                             int32_t nx, int32_t nr, int32_t nkO) {
    CH_TRY(env_check(np, nx, nr, nkO, "proj_fld_env"));
    if (np == 0) return ST_OK;
    const size_t ng = (size_t)(nx + 1) * (nr + 1) * (2 * nkO + 1) * 6;
    Stage st;
    const double* x = st.in(coord, 3 * (size_t)np);
    const double* w = st.in(wghts, (size_t)np);
    const double2* F = reinterpret_cast<const double2*>(st.in(Fld, 2 * ng));
    double* Ft = st.inout(Fld_tot, 6 * (size_t)np);
    const double* R = st.in(Rgrid, (size_t)nr + 1);
    CH_TRY(st.status());
    GridGeom g = make_geom(leftX, R, dx_inv, dr_inv, nx, nr, nkO);
    CH_TRY(load_rgrid_edges(g, Rgrid, R));
    hipLaunchKernelGGL(k_proj_fld_env<double>, blocks_for(np, 256), 256, 0, stream(), x, ENV_AOS3, w, F, Ft,
                       ENV_AOS6, g, kx0, np);
    CH_LAUNCH_CHECK("k_proj_fld_env");
    return st.finish();
}

int32_t chimera_soa_dep_curr_env(const double* coord_half, const double* momenta, const double* wghts, int64_t ld,
                                 int32_t np, double* curr, double leftX, const double* Rgrid, double dx_inv,
                                 double dr_inv, double r0, double rmax, double kx0, int32_t nx, int32_t nr,
                                 int32_t nkO, const double* frame) {
    CH_TRY(env_check(np, nx, nr, nkO, "soa_dep_curr_env"));
    if (ld < np) return fail(ST_EINVAL, "soa_dep_curr_env: ld=%lld < np=%d", (long long)ld, np);
    if (np == 0) return ST_OK;
    if (!is_device_ptr(coord_half) || !is_device_ptr(momenta) || !is_device_ptr(wghts) || !is_device_ptr(curr) ||
        !is_device_ptr(Rgrid))
        return fail(ST_EINVAL, "soa_dep_curr_env: resident entry points take device pointers");
    GridGeom g = resident_geom(leftX, Rgrid, dx_inv, dr_inv, r0, rmax, nx, nr, nkO);
    g.leftX_dev = frame;
    const PStride s = {1, ld};
    KTimer kt(KF_DEP_CURR);
    hipLaunchKernelGGL(k_dep_curr_env, blocks_for(np, 256), 256, 0, stream(), coord_half, s, momenta, s, wghts,
                       reinterpret_cast<double2*>(curr), g, kx0, 0, np);
    CH_LAUNCH_CHECK("k_dep_curr_env");
    return ST_OK;
}

int32_t chimera_soa_dep_dens_env(const double* coord, const double* wghts, int64_t ld, int32_t np, double* dens,
                                 double leftX, const double* Rgrid, double dx_inv, double dr_inv, double r0,
                                 double rmax, double kx0, int32_t nx, int32_t nr, int32_t nkO, const double* frame) {
    CH_TRY(env_check(np, nx, nr, nkO, "soa_dep_dens_env"));
    if (ld < np) return fail(ST_EINVAL, "soa_dep_dens_env: ld=%lld < np=%d", (long long)ld, np);
    if (np == 0) return ST_OK;
    if (!is_device_ptr(coord) || !is_device_ptr(wghts) || !is_device_ptr(dens) || !is_device_ptr(Rgrid))
        return fail(ST_EINVAL, "soa_dep_dens_env: resident entry points take device pointers");
    GridGeom g = resident_geom(leftX, Rgrid, dx_inv, dr_inv, r0, rmax, nx, nr, nkO);
    g.leftX_dev = frame;
    const PStride s = {1, ld};
    KTimer kt(KF_DEP_DENS);
    hipLaunchKernelGGL(k_dep_dens_env<double>, blocks_for(np, 256), 256, 0, stream(), coord, s, wghts,
                       reinterpret_cast<double2*>(dens), g, kx0, 0, np);
    CH_LAUNCH_CHECK("k_dep_dens_env");
    return ST_OK;
}

int32_t chimera_soa_gather_push_env(const double* coord, double* momenta, const double* wghts, int64_t ld, int32_t np,
                                    const double* Fld, double leftX, const double* Rgrid, double dx_inv,
                                    double dr_inv, double r0, double rmax, double kx0, int32_t nx, int32_t nr,
                                    int32_t nkO, double push_dt, const chimera_device_t* devices, int32_t ndev,
                                    double t, const double* frame) {
    CH_TRY(env_check(np, nx, nr, nkO, "soa_gather_push_env"));
    if (ld < np) return fail(ST_EINVAL, "soa_gather_push_env: ld=%lld < np=%d", (long long)ld, np);
    if (np == 0) return ST_OK;
    if (!is_device_ptr(coord) || !is_device_ptr(momenta) || !is_device_ptr(wghts) || !is_device_ptr(Fld) ||
        !is_device_ptr(Rgrid))
        return fail(ST_EINVAL, "soa_gather_push_env: resident entry points take device pointers");
    ExtDevices devs;
    CH_TRY(ext_devices(devs, devices, ndev, t));
    GridGeom g = resident_geom(leftX, Rgrid, dx_inv, dr_inv, r0, rmax, nx, nr, nkO);
    g.leftX_dev = frame;
    const PStride s = {1, ld};
    KTimer kt(KF_GATHER_PUSH);
    hipLaunchKernelGGL(k_gather_push_env<double>, blocks_for(np, 256), 256, 0, stream(), coord, s, momenta, s, wghts,
                       reinterpret_cast<const double2*>(Fld), g, kx0, push_dt, devs, np);
    CH_LAUNCH_CHECK("k_gather_push_env");
    return ST_OK;
}

int32_t chimera_eb_correction_env(double* eb_spc, int32_t nx, int32_t nr, int32_t nkO) {
    if (nx < 0 || nr < 0 || nkO < 0) return fail(ST_EINVAL, "eb_correction_env: bad sizes");
    const int nm = 2 * nkO + 1;
    const size_t ng = (size_t)(nx + 1) * (nr + 1) * nm * 6;
    Stage st;
    double2* E = reinterpret_cast<double2*>(st.inout(eb_spc, 2 * ng));
    CH_TRY(st.status());
    hipLaunchKernelGGL(k_eb_correction_env<double2>, blocks_for((int64_t)(nx + 1) * nm * 6, 256), 256, 0, stream(), E,
                       nx + 1, nr + 1, nm, nkO);
    CH_LAUNCH_CHECK("k_eb_correction_env");
    return st.finish();
}

}  // extern "C"
