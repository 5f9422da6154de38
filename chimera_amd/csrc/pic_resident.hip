// pic_resident.hip -- particle kernels of device-resident simulations.
//
// Particles live in structure-of-arrays form: coordinate component c of
// particle i at coord[c*ld + i] (likewise momenta), weights[i].  Every pointer
// is device memory; calls are asynchronous on the library stream.
//
//  * chimera_soa_push_coords   leap-frog push (particle_tools.f90:58-82)
//  * chimera_soa_dep_curr/dens deposition without the ghost fold, which the
//                              forward FB transform applies when it reads
//                              row 1 (grid_deps.f90:80-85, :144-145)
//  * chimera_soa_gather_push   gather (grid_deps.f90:149-217) + analytic
//                              undulator (devices.f90:162-203) + Boris push
//                              (particle_tools.f90:18-56) in one pass, with
//                              the reference's order: field on particles is
//                              zeroed, gathered, device fields added, pushed
//                              (chimera_main.py:130-151, :91-92)
//  * chimera_soa_compact       stable removal of particles outside lims
//                              (particle_tools.f90:130-153 + the permutation
//                              gathers of species.py:385-398)
#include "common.hpp"
#include "pic_device.hpp"
#include "sort.hpp"
#include "../../include/chimera_hip.h"

namespace chimera {

__global__ void k_gather_push(const double* __restrict__ coord, PStride sx, double* __restrict__ u, PStride su,
                              const double* __restrict__ w, const double2* __restrict__ Fld, GridGeom g,
                              double push_dt, UndulParams up, int n_undul, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    const double xp = coord[sx.at(ip, 0)], yp = coord[sx.at(ip, 1)], zp = coord[sx.at(ip, 2)];
    double f[6] = {0, 0, 0, 0, 0, 0};
    PartGeom pg;
    if (w[ip] != 0.0 && part_geom(pg, xp, yp, zp, g, +1.0)) {
        const int64_t sr = g.nx + 1, sm = sr * (g.nr + 1), sl = sm * (g.nkO + 1);
        const double2* base = Fld + pg.ix + pg.ir * sr;
        for (int l = 0; l < 6; l++) {
            double acc = 0.0;
            for (int iO = 0; iO <= g.nkO; iO++) {
                const double2* fl = base + iO * sm + l * sl;
                for (int b = 0; b < 2; b++)
                    for (int a = 0; a < 2; a++) {
                        const double2 pc = crs(pg.S[1][b] * pg.S[0][a], pg.ph[iO]);
                        const double2 v = fl[a + b * sr];
                        acc += pc.x * v.x - pc.y * v.y;
                    }
            }
            f[l] = acc;
        }
    }
    if (n_undul) undul_field(up, xp, yp, f);
    double m[3] = {u[su.at(ip, 0)], u[su.at(ip, 1)], u[su.at(ip, 2)]};
    boris(m, f, 0.5 * push_dt);
    for (int c = 0; c < 3; c++) u[su.at(ip, c)] = m[c];
}

// Stable compaction: flags, exclusive scan, scatter of every particle array.
__global__ void k_keep_flags(const double* __restrict__ coord, int64_t ld, double l0, double l1, double l2, double l3,
                             int32_t* __restrict__ flag, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    const double x = coord[ip], y = coord[ld + ip], z = coord[2 * ld + ip];
    const double r2 = y * y + z * z;
    flag[ip] = (x >= l0 && x <= l1 && r2 >= l2 && r2 <= l3) ? 1 : 0;
}

__global__ void k_compact_scatter(const int32_t* __restrict__ flag, const int64_t* __restrict__ pos,
                                  const double* __restrict__ src, double* __restrict__ dst, int64_t ld_src,
                                  int64_t ld_dst, int ncomp, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np || !flag[ip]) return;
    const int64_t o = pos[ip];
    for (int c = 0; c < ncomp; c++) dst[c * ld_dst + o] = src[c * ld_src + ip];
}

static GridGeom resident_geom(double leftX, const double* Rgrid, double dx_inv, double dr_inv, double r0, double rmax,
                              int nx, int nr, int nkO) {
    GridGeom g = make_geom(leftX, Rgrid, dx_inv, dr_inv, nx, nr, nkO);
    g.r0 = r0;
    g.rmax = rmax;
    return g;
}

static int32_t check_resident(int np, int64_t ld, int nkO) {
    if (np < 0 || ld < np) return fail(ST_EINVAL, "resident particles: np=%d ld=%lld", np, (long long)ld);
    if (nkO < 0 || nkO >= MAX_MODES)
        return fail(ST_EUNSUP, "azimuthal modes 0..%d requested; this build supports up to %d modes", nkO, MAX_MODES);
    return ST_OK;
}

}  // namespace chimera

using namespace chimera;

extern "C" {

int32_t chimera_soa_push_coords(double* coord, double* coord_half, const double* momenta, int64_t ld, int32_t np,
                                double dt) {
    CH_TRY(check_resident(np, ld, 0));
    if (np == 0) return ST_OK;
    const PStride s = {1, ld};
    return launch_push_coords(coord, s, momenta, s, coord_half, s, dt, np);
}

int32_t chimera_soa_dep_curr(const double* coord, const double* momenta, const double* wghts, int64_t ld, int32_t np,
                             double* curr, double leftX, const double* Rgrid, double dx_inv, double dr_inv, double r0,
                             double rmax, int32_t nx, int32_t nr, int32_t nkO) {
    CH_TRY(check_resident(np, ld, nkO));
    if (np == 0) return ST_OK;
    const PStride s = {1, ld};
    const GridGeom g = resident_geom(leftX, Rgrid, dx_inv, dr_inv, r0, rmax, nx, nr, nkO);
    return launch_dep_curr(coord, s, momenta, s, wghts, reinterpret_cast<double2*>(curr), g, 0, np);
}

int32_t chimera_soa_dep_dens(const double* coord, const double* wghts, int64_t ld, int32_t np, double* dens,
                             double leftX, const double* Rgrid, double dx_inv, double dr_inv, double r0, double rmax,
                             int32_t nx, int32_t nr, int32_t nkO) {
    CH_TRY(check_resident(np, ld, nkO));
    if (np == 0) return ST_OK;
    const PStride s = {1, ld};
    const GridGeom g = resident_geom(leftX, Rgrid, dx_inv, dr_inv, r0, rmax, nx, nr, nkO);
    return launch_dep_dens(coord, s, wghts, reinterpret_cast<double2*>(dens), g, 0, np);
}

int32_t chimera_soa_gather_push(const double* coord, double* momenta, const double* wghts, int64_t ld, int32_t np,
                                const double* Fld, double leftX, const double* Rgrid, double dx_inv, double dr_inv,
                                double r0, double rmax, int32_t nx, int32_t nr, int32_t nkO, double push_dt,
                                const double* undul_params) {
    CH_TRY(check_resident(np, ld, nkO));
    if (np == 0) return ST_OK;
    const PStride s = {1, ld};
    const GridGeom g = resident_geom(leftX, Rgrid, dx_inv, dr_inv, r0, rmax, nx, nr, nkO);
    UndulParams up = {0, 1, 0, 0, 0};
    int nu = 0;
    if (undul_params) {
        up = make_undul(undul_params);
        nu = 1;
    }
    hipLaunchKernelGGL(k_gather_push, blocks_for(np, 256), 256, 0, stream(), coord, s, momenta, s, wghts,
                       reinterpret_cast<const double2*>(Fld), g, push_dt, up, nu, np);
    CH_LAUNCH_CHECK("k_gather_push");
    return ST_OK;
}

int32_t chimera_soa_compact(double* coord, double* coord_half, double* momenta, double* wghts, int64_t ld, int32_t np,
                            const double* lims, int32_t test_half, int32_t* np_out) {
    CH_TRY(check_resident(np, ld, 0));
    if (lims == nullptr || np_out == nullptr) return fail(ST_EINVAL, "soa_compact: null lims/np_out");
    if (np == 0) {
        *np_out = 0;
        return ST_OK;
    }
    int32_t* flag = static_cast<int32_t*>(scratch(4, (size_t)np * sizeof(int32_t)));
    int64_t* pos = static_cast<int64_t*>(scratch(5, ((size_t)np + 8) * sizeof(int64_t)));
    double* tmp = static_cast<double*>(scratch(7, (size_t)np * 3 * sizeof(double)));
    if (!flag || !pos || !tmp) return ST_ENOMEM;
    int64_t* total = pos + np;
    const double* test = test_half ? coord_half : coord;
    hipLaunchKernelGGL(k_keep_flags, blocks_for(np, 256), 256, 0, stream(), test, ld, lims[0], lims[1], lims[2],
                       lims[3], flag, np);
    CH_LAUNCH_CHECK("k_keep_flags");
    CH_TRY(exclusive_scan(flag, pos, np, total, 6));
    double* arrays[4] = {coord, coord_half, momenta, wghts};
    const int ncomp[4] = {3, 3, 3, 1};
    for (int a = 0; a < 4; a++) {
        hipLaunchKernelGGL(k_compact_scatter, blocks_for(np, 256), 256, 0, stream(), flag, pos, arrays[a], tmp, ld,
                           (int64_t)np, ncomp[a], np);
        CH_LAUNCH_CHECK("k_compact_scatter");
        CH_HIP(hipMemcpy2DAsync(arrays[a], ld * sizeof(double), tmp, (size_t)np * sizeof(double),
                                (size_t)np * sizeof(double), ncomp[a], hipMemcpyDeviceToDevice, stream()));
    }
    int64_t h = 0;
    CH_HIP(hipMemcpyAsync(&h, total, sizeof h, hipMemcpyDeviceToHost, stream()));
    CH_HIP(hipStreamSynchronize(stream()));
    *np_out = (int32_t)h;
    return ST_OK;
}

}  // extern "C"
