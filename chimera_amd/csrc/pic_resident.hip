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
#include <climits>
#include <cstdlib>

#include "common.hpp"
#include "pic_device.hpp"
#include "sort.hpp"
#include "../../include/chimera_hip.h"

namespace chimera {

__global__ void k_gather_push(const double* __restrict__ coord, PStride sx, double* __restrict__ u, PStride su,
                              const double* __restrict__ w, const double2* __restrict__ Fld, GridGeom g,
                              double push_dt, UndulParams up, int n_undul, int np) {
    const int ip = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
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

// ------------------------------------------------------------------ fused push + deposition
//
// One block takes DEP_PPT consecutive particles per thread (thread t owns
// particles p0 + t*DEP_PPT + j, so the lanes of a wave sit DEP_PPT particles
// apart and, with particles in cell order, mostly in different cells).
// Phase 1 optionally pushes positions (particle_tools.f90:58-82) and finds the
// block's cell bounding box for J (at the half-step positions) and rho (at the
// new positions).  Phase 2 accumulates into LDS tiles covering those boxes with
// LDS FP64 atomics; phase 3 adds the tiles to the grid with one global atomic
// per node and component.  A box too large for the LDS budget makes that block
// deposit straight to the grid with global atomics.  Deposits are unfolded; the
// forward FB transform applies the ghost fold.


struct DepArgs {
    double* x;          // positions (3, ld), updated when push
    double* xh;         // half-step positions (3, ld): written when push, read otherwise
    const double* u;
    const double* w;
    int64_t ld;
    int np;
    int push;
    double dt;
    GridGeom gJ, gR;    // geometry for J (leftX at the current deposit) and rho
    double2* J;         // (nx+1, nr+1, M, 3) or null
    double2* R;         // (nx+1, nr+1, M) or null
};

__device__ __forceinline__ void lds_add(double* p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ bool cell_of(double xp, double yp, double zp, const GridGeom& g, int& ix, int& ir) {
    const double rp = sqrt(yp * yp + zp * zp);
    if (!(rp < g.rmax)) return false;
    const double fx = (xp - g.leftX) * g.dx_inv, fr = (rp - g.r0) * g.dr_inv;
    if (!(fx >= 0.0 && fx < (double)g.nx) || !(fr >= 0.0 && fr < (double)g.nr)) return false;
    ix = (int)floor(fx);
    ir = (int)floor(fr);
    return true;
}

__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

template <int DEP_THREADS, int DEP_PPT, int DEP_LDS_BYTES>
__global__ void __launch_bounds__(DEP_THREADS) k_push_deposit(DepArgs a) {
    constexpr int DEP_CHUNK = DEP_THREADS * DEP_PPT;
    extern __shared__ double tile[];
    __shared__ int red[8][DEP_THREADS / 64];
    __shared__ int box[8];
    const int tid = threadIdx.x;
    const int64_t p0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * DEP_CHUNK;
    const int nmod = a.gJ.nkO + 1;
    const int64_t ld = a.ld;
    const bool doJ = a.J != nullptr, doR = a.R != nullptr;

    // ---- phase 1: push, bounding boxes
    int bx[8] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN};
    for (int j = 0; j < DEP_PPT; j++) {
        const int64_t ip = p0 + (int64_t)tid * DEP_PPT + j;
        if (ip >= a.np) break;
        const double wp = a.w[ip];
        double xo[3] = {a.x[ip], a.x[ld + ip], a.x[2 * ld + ip]};
        double xn[3], xhh[3];
        const double u0 = a.u[ip], u1 = a.u[ld + ip], u2 = a.u[2 * ld + ip];
        if (a.push) {
            const double dt_gp = a.dt / sqrt(1.0 + (u0 * u0 + u1 * u1 + u2 * u2));
            const double uu[3] = {u0, u1, u2};
            for (int c = 0; c < 3; c++) {
                xn[c] = xo[c] + uu[c] * dt_gp;
                xhh[c] = 0.5 * (xo[c] + xn[c]);
                a.x[c * ld + ip] = xn[c];
                a.xh[c * ld + ip] = xhh[c];
            }
        } else {
            for (int c = 0; c < 3; c++) {
                xn[c] = xo[c];
                xhh[c] = a.xh[c * ld + ip];
            }
        }
        if (wp == 0.0) continue;
        int ix, ir;
        if (doJ && (fabs(u0) + fabs(u1) + fabs(u2)) != 0.0 && cell_of(xhh[0], xhh[1], xhh[2], a.gJ, ix, ir)) {
            bx[0] = min(bx[0], ix); bx[1] = max(bx[1], ix); bx[2] = min(bx[2], ir); bx[3] = max(bx[3], ir);
        }
        if (doR && cell_of(xn[0], xn[1], xn[2], a.gR, ix, ir)) {
            bx[4] = min(bx[4], ix); bx[5] = max(bx[5], ix); bx[6] = min(bx[6], ir); bx[7] = max(bx[7], ir);
        }
    }
    for (int q = 0; q < 8; q++) {
        const int v = (q & 1) ? wave_max(bx[q]) : wave_min(bx[q]);
        if ((tid & 63) == 0) red[q][tid >> 6] = v;
    }
    __syncthreads();
    if (tid < 8) {
        int v = red[tid][0];
        for (int wv = 1; wv < DEP_THREADS / 64; wv++) v = (tid & 1) ? max(v, red[tid][wv]) : min(v, red[tid][wv]);
        box[tid] = v;
    }
    __syncthreads();
    // tiles: J nodes (TXJ x TRJ) x 3 comps x nmod modes, then rho nodes x nmod; complex = 2 doubles
    const int TXJ = box[1] >= box[0] ? box[1] - box[0] + 2 : 0, TRJ = box[3] >= box[2] ? box[3] - box[2] + 2 : 0;
    const int TXR = box[5] >= box[4] ? box[5] - box[4] + 2 : 0, TRR = box[7] >= box[6] ? box[7] - box[6] + 2 : 0;
    const int64_t nJ = (int64_t)TXJ * TRJ * 3 * nmod * 2, nR = (int64_t)TXR * TRR * nmod * 2;
    const bool tileJ = doJ && nJ > 0 && nJ * 8 <= DEP_LDS_BYTES;
    const bool tileR = doR && nR > 0 && (nR + (tileJ ? nJ : 0)) * 8 <= DEP_LDS_BYTES;
    double* tJ = tile;
    double* tR = tile + (tileJ ? nJ : 0);
    const int64_t nz = (tileJ ? nJ : 0) + (tileR ? nR : 0);
    for (int64_t i = tid; i < nz; i += DEP_THREADS) tile[i] = 0.0;
    __syncthreads();

    // ---- phase 2: deposit
    const int64_t sr = a.gJ.nx + 1, sm = sr * (a.gJ.nr + 1), sl = sm * nmod;
    for (int j = 0; j < DEP_PPT; j++) {
        const int64_t ip = p0 + (int64_t)tid * DEP_PPT + j;
        if (ip >= a.np) break;
        const double wp = a.w[ip];
        if (wp == 0.0) continue;
        PartGeom pg;
        if (doJ) {
            double v[3] = {a.u[ip], a.u[ld + ip], a.u[2 * ld + ip]};
            if ((fabs(v[0]) + fabs(v[1]) + fabs(v[2])) != 0.0 &&
                part_geom(pg, a.xh[ip], a.xh[ld + ip], a.xh[2 * ld + ip], a.gJ, -1.0)) {
                const double gp = sqrt(1 + v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                for (int c = 0; c < 3; c++) v[c] = v[c] * wp / gp;
                // in-tile test on the cell found here (phase 1 may round differently at a cell edge)
                const int64_t tx = pg.ix - box[0], tr = pg.ir - box[2];
                const bool inJ = tileJ && tx >= 0 && tx + 1 < TXJ && tr >= 0 && tr + 1 < TRJ;
                for (int l = 0; l < 3; l++)
                    for (int m = 0; m < nmod; m++) {
                        const double2 pv = crs(v[l], pg.ph[m]);
                        for (int b = 0; b < 2; b++)
                            for (int aa = 0; aa < 2; aa++) {
                                const double2 d = crs(pg.S[0][aa] * pg.S[1][b], pv);
                                if (inJ) {
                                    const int64_t t = (((int64_t)(l * nmod + m) * TRJ + (tr + b)) * TXJ + (tx + aa)) * 2;
                                    lds_add(tJ + t, d.x);
                                    lds_add(tJ + t + 1, d.y);
                                } else {
                                    double* dst = reinterpret_cast<double*>(a.J + (pg.ix + aa) + (pg.ir + b) * sr +
                                                                            m * sm + l * sl);
                                    unsafeAtomicAdd(dst, d.x);
                                    unsafeAtomicAdd(dst + 1, d.y);
                                }
                            }
                    }
            }
        }
        if (doR && part_geom(pg, a.x[ip], a.x[ld + ip], a.x[2 * ld + ip], a.gR, -1.0)) {
            const int64_t tx = pg.ix - box[4], tr = pg.ir - box[6];
            const bool inR = tileR && tx >= 0 && tx + 1 < TXR && tr >= 0 && tr + 1 < TRR;
            for (int m = 0; m < nmod; m++)
                for (int b = 0; b < 2; b++)
                    for (int aa = 0; aa < 2; aa++) {
                        const double2 d = crs(pg.S[0][aa] * pg.S[1][b] * wp, pg.ph[m]);
                        if (inR) {
                            const int64_t t = (((int64_t)m * TRR + (tr + b)) * TXR + (tx + aa)) * 2;
                            lds_add(tR + t, d.x);
                            lds_add(tR + t + 1, d.y);
                        } else {
                            double* dst = reinterpret_cast<double*>(a.R + (pg.ix + aa) + (pg.ir + b) * sr + m * sm);
                            unsafeAtomicAdd(dst, d.x);
                            unsafeAtomicAdd(dst + 1, d.y);
                        }
                    }
        }
    }
    __syncthreads();

    // ---- phase 3: flush the tiles (x fastest: coalesced atomics)
    if (tileJ) {
        const int64_t n = nJ / 2;
        for (int64_t i = tid; i < n; i += DEP_THREADS) {
            const double re = tJ[2 * i], im = tJ[2 * i + 1];
            if (re == 0.0 && im == 0.0) continue;
            const int64_t xx = i % TXJ, rest = i / TXJ, rr = rest % TRJ, lm = rest / TRJ;
            double* dst = reinterpret_cast<double*>(a.J + (box[0] + xx) + (box[2] + rr) * sr + lm * sm);
            unsafeAtomicAdd(dst, re);
            unsafeAtomicAdd(dst + 1, im);
        }
    }
    if (tileR) {
        const int64_t n = nR / 2;
        for (int64_t i = tid; i < n; i += DEP_THREADS) {
            const double re = tR[2 * i], im = tR[2 * i + 1];
            if (re == 0.0 && im == 0.0) continue;
            const int64_t xx = i % TXR, rest = i / TXR, rr = rest % TRR, m = rest / TRR;
            double* dst = reinterpret_cast<double*>(a.R + (box[4] + xx) + (box[6] + rr) * sr + m * sm);
            unsafeAtomicAdd(dst, re);
            unsafeAtomicAdd(dst + 1, im);
        }
    }
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

// ------------------------------------------------------------------ counting sort by cell
//
// key = ir*nx + ix for a particle in cell (ix, ir) of the grid, nx*nr for a
// kept particle outside the grid, nx*nr + 1 for a particle outside lims (it is
// dropped).  The cell comes from the same expressions as grid_deps.f90:46-47,
// evaluated with explicitly rounded operations so that the integer key of a
// particle does not depend on contraction choices of the compiler.
__global__ void k_cell_keys(const double* __restrict__ x, const double* __restrict__ xh, int64_t ld, int np,
                            double leftX, double dx_inv, double dr_inv, double r0, int nx, int nr, int use_lims,
                            double l0, double l1, double l2, double l3, int test_half, uint32_t* __restrict__ keys,
                            int32_t* __restrict__ kept) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = ip < np;
    uint32_t key = 0;
    if (valid) {
        const double px = x[ip], py = x[ld + ip], pz = x[2 * ld + ip];
        const uint32_t ncell = (uint32_t)nx * (uint32_t)nr;
        bool drop = false;
        if (use_lims) {
            const double* t = test_half ? xh : x;
            const double tx = t[ip], ty = t[ld + ip], tz = t[2 * ld + ip];
            const double r2 = __dadd_rn(__dmul_rn(ty, ty), __dmul_rn(tz, tz));
            drop = !(tx >= l0 && tx <= l1 && r2 >= l2 && r2 <= l3);
        }
        if (drop) {
            key = ncell + 1;
        } else {
            const double rp = sqrt(__dadd_rn(__dmul_rn(py, py), __dmul_rn(pz, pz)));
            const double fx = __dmul_rn(__dsub_rn(px, leftX), dx_inv);
            const double fr = __dmul_rn(__dsub_rn(rp, r0), dr_inv);
            if (fx >= 0.0 && fx < (double)nx && fr >= 0.0 && fr < (double)nr)
                key = (uint32_t)floor(fr) * (uint32_t)nx + (uint32_t)floor(fx);
            else
                key = ncell;
        }
        keys[ip] = key;
    }
    // kept-particle count, one atomic per wave
    const uint64_t m = __ballot(valid && key <= (uint32_t)nx * (uint32_t)nr);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(kept, (int32_t)__popcll(m));
}

__global__ void k_gather_particles(const int32_t* __restrict__ idx, int n, const double* __restrict__ x,
                                   const double* __restrict__ xh, const double* __restrict__ u,
                                   const double* __restrict__ w, int64_t ld, double* __restrict__ x2,
                                   double* __restrict__ xh2, double* __restrict__ u2, double* __restrict__ w2,
                                   int64_t ld2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t j = idx[i];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        x2[c * ld2 + i] = x[c * ld + j];
        xh2[c * ld2 + i] = xh[c * ld + j];
        u2[c * ld2 + i] = u[c * ld + j];
    }
    w2[i] = w[j];
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
    KTimer kt(KF_GATHER_PUSH);
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
    KTimer kt(KF_SORT);
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


int32_t chimera_soa_push_deposit(double* coord, double* coord_half, const double* momenta, const double* wghts,
                                 int64_t ld, int32_t np, double dt, double leftX_J, double leftX_rho,
                                 const double* Rgrid, double dx_inv, double dr_inv, double r0, double rmax, int32_t nx,
                                 int32_t nr, int32_t nkO, double* curr, double* dens) {
    CH_TRY(check_resident(np, ld, nkO));
    if (np == 0) return ST_OK;
    DepArgs a;
    a.x = coord;
    a.xh = coord_half;
    a.u = momenta;
    a.w = wghts;
    a.ld = ld;
    a.np = np;
    a.push = dt != 0.0;
    a.dt = dt;
    a.gJ = resident_geom(leftX_J, Rgrid, dx_inv, dr_inv, r0, rmax, nx, nr, nkO);
    a.gR = resident_geom(leftX_rho, Rgrid, dx_inv, dr_inv, r0, rmax, nx, nr, nkO);
    a.J = reinterpret_cast<double2*>(curr);
    a.R = reinterpret_cast<double2*>(dens);
    // block shape: threads x consecutive particles per thread, LDS tile budget;
    // CHIMERA_DEP_SHAPE selects one for sweeps
    static int shape = -1;
    if (shape < 0) {
        const char* e = getenv("CHIMERA_DEP_SHAPE");
        shape = e ? atoi(e) : 4;
        if (shape < 0 || shape > 4) shape = 4;
    }
    KTimer kt(a.J ? KF_DEP_CURR : KF_DEP_DENS);
    switch (shape) {
        case 0:
            hipLaunchKernelGGL((k_push_deposit<256, 4, 48 * 1024>), blocks_for(np, 1024), 256, 48 * 1024, stream(), a);
            break;
        case 1:
            hipLaunchKernelGGL((k_push_deposit<64, 16, 48 * 1024>), blocks_for(np, 1024), 64, 48 * 1024, stream(), a);
            break;
        case 2:
            hipLaunchKernelGGL((k_push_deposit<64, 8, 24 * 1024>), blocks_for(np, 512), 64, 24 * 1024, stream(), a);
            break;
        case 3:
            hipLaunchKernelGGL((k_push_deposit<128, 4, 24 * 1024>), blocks_for(np, 512), 128, 24 * 1024, stream(), a);
            break;
        default:
            hipLaunchKernelGGL((k_push_deposit<64, 4, 12 * 1024>), blocks_for(np, 256), 64, 12 * 1024, stream(), a);
            break;
    }
    CH_LAUNCH_CHECK("k_push_deposit");
    return ST_OK;
}


int32_t chimera_soa_sort_cells(const double* coord, const double* coord_half, const double* momenta,
                               const double* wghts, int64_t ld, int32_t np, double* coord2, double* coord_half2,
                               double* momenta2, double* wghts2, int64_t ld2, double leftX, double dx_inv,
                               double dr_inv, double r0, int32_t nx, int32_t nr, const double* lims, int32_t test_half,
                               int32_t* np_out, int32_t* perm) {
    CH_TRY(check_resident(np, ld, 0));
    if (np_out == nullptr || ld2 < np || nx < 1 || nr < 1) return fail(ST_EINVAL, "soa_sort_cells: bad arguments");
    if ((int64_t)nx * nr + 2 > (int64_t)UINT32_MAX) return fail(ST_EINVAL, "soa_sort_cells: too many cells");
    if (np == 0) {
        *np_out = 0;
        return ST_OK;
    }
    KTimer kt(KF_SORT);
    uint32_t* keys = static_cast<uint32_t*>(scratch(8, (size_t)np * sizeof(uint32_t)));
    int32_t* idx = perm ? perm : static_cast<int32_t*>(scratch(9, (size_t)np * sizeof(int32_t)));
    uint32_t* ktmp = static_cast<uint32_t*>(scratch(10, (size_t)np * sizeof(uint32_t)));
    int32_t* itmp = static_cast<int32_t*>(scratch(11, (size_t)np * sizeof(int32_t)));
    int32_t* kept = static_cast<int32_t*>(scratch(15, 256));
    if (!keys || !idx || !ktmp || !itmp || !kept) return ST_ENOMEM;
    CH_HIP(hipMemsetAsync(kept, 0, sizeof(int32_t), stream()));
    double L[4] = {0, 0, 0, 0};
    if (lims) memcpy(L, lims, sizeof L);
    hipLaunchKernelGGL(k_cell_keys, blocks_for(np, 256), 256, 0, stream(), coord, coord_half, ld, np, leftX, dx_inv,
                       dr_inv, r0, nx, nr, lims ? 1 : 0, L[0], L[1], L[2], L[3], test_half, keys, kept);
    CH_LAUNCH_CHECK("k_cell_keys");
    const uint64_t maxkey = (uint64_t)nx * nr + 1;
    int bits = 1;
    while ((1ull << bits) <= maxkey) bits++;
    CH_TRY(radix_sort_keys(keys, idx, ktmp, itmp, np, bits, 12, 13));
    int32_t hk = 0;
    CH_HIP(hipMemcpyAsync(&hk, kept, sizeof hk, hipMemcpyDeviceToHost, stream()));
    CH_HIP(hipStreamSynchronize(stream()));
    hipLaunchKernelGGL(k_gather_particles, blocks_for(hk, 256), 256, 0, stream(), idx, hk, coord, coord_half, momenta,
                       wghts, ld, coord2, coord_half2, momenta2, wghts2, ld2);
    CH_LAUNCH_CHECK("k_gather_particles");
    *np_out = hk;
    return ST_OK;
}

}  // extern "C"
