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
#include <type_traits>

#include "common.hpp"
#include "fb_kernels.hpp"
#include "pic_device.hpp"
#include "sort.hpp"
#include "../../include/chimera_hip.h"

namespace chimera {

// NMOD = number of azimuthal modes, a template parameter so that every mode
// index is a constant and the per-particle phases stay in registers.
// R: precision of momenta, weights and fields (double, or float in mixed
// precision, where the Boris push runs in FP32); positions and the cell
// geometry are FP64.
//
// A block takes GP_T consecutive (cell-ordered) particles, one per thread,
// finds the bounding box of their cells, and copies the field nodes of that
// box (all six components and modes) into LDS with coalesced loads; each
// particle then reads its 4 x 6 x NMOD values from LDS instead of from L2 (at
// 16 particles per cell the block's 256 particles share 17 x 3 nodes).  The
// momenta load with the positions, in flight while the tile loads.  A box
// too large for GP_LDS_BYTES reads the field from global memory.  The values
// and the order of the sums are the same either way.
constexpr int GP_T = 256, GP_LDS_BYTES = 24 * 1024;

__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int block_min_max(int v, bool is_max, int (*red)[GP_T / 64], int q) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    v = is_max ? wave_max(v) : wave_min(v);
    if (lane == 0) red[q][wv] = v;
    return v;
}

template <int NMOD, typename R>
__global__ void __launch_bounds__(GP_T) k_gather_push(const double* __restrict__ coord, PStride sx, R* __restrict__ u,
                                                      PStride su, const R* __restrict__ w,
                                                      const cplx_t<R>* __restrict__ Fld, GridGeom g_in, double push_dt,
                                                      ExtDevices devs, int np) {
    typedef cplx_t<R> C;
    constexpr int NC = 6 * NMOD;   // (component, mode) planes
    __shared__ C tile[GP_LDS_BYTES / sizeof(C)];
    __shared__ int red[4][GP_T / 64];
    __shared__ int box[4];
    const int tid = threadIdx.x;
    const int ip = xcd_block(blockIdx.x, gridDim.x) * GP_T + tid;
    const bool live = ip < np;
    const GridGeom g = frame_geom(g_in);
    double xp = 0, yp = 0, zp = 0;
    R m[3] = {0, 0, 0};
    if (live) {
        xp = coord[sx.at(ip, 0)];
        yp = coord[sx.at(ip, 1)];
        zp = coord[sx.at(ip, 2)];
        // the momenta load now, so that it is in flight while the field tile loads
        m[0] = u[su.at(ip, 0)];
        m[1] = u[su.at(ip, 1)];
        m[2] = u[su.at(ip, 2)];
    }
    PartGeom pg;
    const bool has = live && w[ip] != 0 && part_geom(pg, xp, yp, zp, g, +1.0);
    block_min_max(has ? (int)pg.ix : INT_MAX, false, red, 0);
    block_min_max(has ? (int)pg.ix : INT_MIN, true, red, 1);
    block_min_max(has ? (int)pg.ir : INT_MAX, false, red, 2);
    block_min_max(has ? (int)pg.ir : INT_MIN, true, red, 3);
    __syncthreads();
    if (tid < 4) {
        int v = red[tid][0];
        for (int k = 1; k < GP_T / 64; k++) v = (tid & 1) ? max(v, red[tid][k]) : min(v, red[tid][k]);
        box[tid] = v;
    }
    __syncthreads();
    const int64_t sr = g.nx + 1, sm = sr * (g.nr + 1), sl = sm * NMOD;
    const int bx0 = box[0], br0 = box[2];
    const int TX = box[1] >= box[0] ? box[1] - box[0] + 2 : 0, TR = box[3] >= box[2] ? box[3] - box[2] + 2 : 0;
    const int nt = TX * TR * NC;
    const bool tiled = nt > 0 && (int64_t)nt * (int64_t)sizeof(C) <= GP_LDS_BYTES;
    if (tiled) {
        // plane lm = l*NMOD + m of the field starts at lm*sm (sl = NMOD*sm)
        for (int e = tid; e < nt; e += GP_T) {
            const int xx = e % TX, rest = e / TX, rr = rest % TR, lm = rest / TR;
            tile[e] = Fld[(bx0 + xx) + (int64_t)(br0 + rr) * sr + (int64_t)lm * sm];
        }
    }
    __syncthreads();
    if (!live) return;
    R f[6] = {0, 0, 0, 0, 0, 0};
    if (has) {
        if (tiled) {
            const C* base = tile + (pg.ix - bx0) + (pg.ir - br0) * TX;
            const int tsl = TX * TR;
            for (int l = 0; l < 6; l++) {
                R acc = 0;
#pragma unroll
                for (int iO = 0; iO < NMOD; iO++) {
                    const C* fl = base + (l * NMOD + iO) * tsl;
                    for (int b = 0; b < 2; b++)
                        for (int a = 0; a < 2; a++) {
                            const C pc = cconv<R>(crs(pg.S[1][b] * pg.S[0][a], pg.ph[iO]));
                            const C v = fl[a + b * TX];
                            acc += pc.x * v.x - pc.y * v.y;
                        }
                }
                f[l] = acc;
            }
        } else {
            const C* base = Fld + pg.ix + pg.ir * sr;
            for (int l = 0; l < 6; l++) {
                R acc = 0;
#pragma unroll
                for (int iO = 0; iO < NMOD; iO++) {
                    const C* fl = base + iO * sm + l * sl;
                    for (int b = 0; b < 2; b++)
                        for (int a = 0; a < 2; a++) {
                            const C pc = cconv<R>(crs(pg.S[1][b] * pg.S[0][a], pg.ph[iO]));
                            const C v = fl[a + b * sr];
                            acc += pc.x * v.x - pc.y * v.y;
                        }
                }
                f[l] = acc;
            }
        }
    }
    if (devs.n) ext_fields<R>(devs, xp, yp, zp, f);
    boris<R>(m, f, (R)(0.5 * push_dt));
    for (int c = 0; c < 3; c++) u[su.at(ip, c)] = m[c];
}

// ------------------------------------------------------------------ fused push + deposition
//
// One block takes PPT consecutive particles per thread (thread t owns
// particles p0 + t*PPT + j, so the lanes of a wave sit PPT particles
// apart and, with particles in cell order, mostly in different cells).
// Phase 1 optionally pushes positions (particle_tools.f90:58-82) and finds the
// block's cell bounding box for J (at the half-step positions) and rho (at the
// new positions).  Then, one component (J_x, J_y, J_z, rho) at a time, the
// block accumulates into an LDS tile covering the box with LDS FP64 atomics
// and hands the tile on:
//  * a block whose tile spans at most DEP_SLOT_R node rows and 8*PPT node
//    columns (the common case for particles in the reference's generation
//    order or sorted by cell: 512 particles cover about 32 cells of one or two
//    cell rows) stores its tile with plain stores in a slot of its own and
//    records (first node row, first node column, width, rows) in a
//    descriptor; k_dep_combine then adds the slots into the grid, so no
//    global atomics;
//  * any other block adds its tile to the grid with one global atomic per
//    node and component, and a box too large for the LDS budget makes the
//    block deposit straight to the grid with global atomics, one per node
//    value and cell from the same per-cell register sums (a block that spans
//    the end of one cell row and the start of the next is such a block).
// Deposits are unfolded; the forward FB transform applies the ghost fold.


template <typename RT>
struct DepArgsT {
    typedef cplx_t<RT> C;
    double* x;          // positions (3, ld), updated when push
    double* xh;         // half-step positions (3, ld): written when push, read otherwise
    const RT* u;        // momenta (3, ld)
    const RT* w;
    int64_t ld;
    int np;
    int push;
    double dt;
    GridGeom gJ, gR;    // geometry for J (leftX at the current deposit) and rho
    C* J;               // (nx+1, nr+1, M, 3) or null
    C* R;               // (nx+1, nr+1, M) or null
    C* slotJ;           // [block][3*M][DEP_SLOT_R][8*PPT] (with J)
    C* slotR;           // [block][M][DEP_SLOT_R][8*PPT] (with R)
    int4* desc;         // [2][blocks]: (first node row or -1, first node column, width, rows) of J, then of R
    int nblocks;
    int vec4;           // particle arrays allow 32-byte vector access (ld % 4 == 0, 32-byte aligned bases)
};

typedef double dv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ dv4 ld4(const double* p) { return *reinterpret_cast<const dv4*>(p); }
__device__ __forceinline__ fv4 ld4(const float* p) { return *reinterpret_cast<const fv4*>(p); }
__device__ __forceinline__ void st4(double* p, dv4 v) { *reinterpret_cast<dv4*>(p) = v; }

// A thread owns PPT consecutive particles, handled in groups of DEP_G (one
// 32-byte vector load per component and group).
constexpr int DEP_THREADS = 64, DEP_G = 4;
// LDS tile of a block: one component (J_x, J_y, J_z or rho) at a time, all
// modes, complex FP64: 16 KiB hold 1024 / nmod nodes -- the 33 x 3 nodes of 512
// cell-ordered particles at 16 per cell up to nmod = 4 -- and eight blocks per
// CU use 128 of the 160 KiB of LDS.  Eight particles per thread (512 per
// block) with 64-column slots: at lpa_cfg2 on cell-sorted particles the
// deposit takes 6.9 ms per FP64 step against 7.5 ms with four per thread and
// 32-column slots (with those slots, eight per thread had overflowed them into
// global atomics and measured slower).
constexpr int DEP_LDS_BYTES = 16 * 1024;
constexpr int DEP_SLOT_R = 4;   // node rows of one slot; its columns: 8 per particle of a thread
// Particles per thread (PPT) by launch: 8 (512 per block, 64-column slots) when
// the particles average at least 8 per cell, else 4 (256 per block, 32-column
// slots): a sparse bunch's 512 particles span more columns than a slot holds.

__device__ __forceinline__ void lds_add(double* p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(float* p, float v) {
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


// R: precision of momenta, weights and grids (double, or float in mixed
// precision).  Positions, the push and the cell geometry stay FP64; each
// corner contribution is formed and summed per cell in FP64, and the LDS tile
// is FP64 in both precisions (FP64 LDS atomics measured far cheaper here than
// FP32 ones under this access pattern), rounded to R when it leaves the block.
template <int NMOD, typename R, int PPT>
__global__ void __launch_bounds__(DEP_THREADS) k_push_deposit(DepArgsT<R> a) {
    constexpr int SW = 8 * PPT;   // slot columns
    typedef cplx_t<R> C;
    constexpr int DEP_CHUNK = DEP_THREADS * PPT;
    __shared__ double tile[DEP_LDS_BYTES / sizeof(double)];
    __shared__ int red[8][DEP_THREADS / 64];
    __shared__ int box[8];
    const int tid = threadIdx.x;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const GridGeom gJ = frame_geom(a.gJ), gR = frame_geom(a.gR);
    const int64_t p0 = (int64_t)blk * DEP_CHUNK;
    constexpr int nmod = NMOD;
    const int64_t ld = a.ld;
    const bool doJ = a.J != nullptr, doR = a.R != nullptr;

    // ---- phase 1: load the thread's PPT consecutive particles, DEP_G at a
    // time (32-byte vector loads when the arrays allow: with PPT = DEP_G a
    // wave then reads 2 KB contiguous per component), push, store, bounding
    // boxes.
    const int64_t ib0 = p0 + (int64_t)tid * PPT;
    // group h of the thread's particles: component c of array p at particle ib0 + h*DEP_G + j
    auto load_parts = [&](const auto* p, double (&v)[3][DEP_G], int ncomp, int h) __attribute__((always_inline)) {
        const int64_t ib = ib0 + h * DEP_G;
        if (a.vec4 && ib + DEP_G <= a.np) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                if (c >= ncomp) break;
                const auto t = ld4(p + c * ld + ib);
                v[c][0] = t.x, v[c][1] = t.y, v[c][2] = t.z, v[c][3] = t.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < DEP_G; j++)
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    if (c >= ncomp) break;
                    v[c][j] = ib + j < a.np ? p[c * ld + ib + j] : 0.0;
                }
        }
    };
    int bx[8] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN};
    for (int h = 0; h < PPT / DEP_G; h++) {
    const int64_t ib = ib0 + h * DEP_G;
    const bool full = a.vec4 && ib + DEP_G <= a.np;
    double X[3][DEP_G], XH[3][DEP_G], U[3][DEP_G], W[3][DEP_G];   // W: row 0 only
    load_parts(a.x, X, 3, h);
    load_parts(a.u, U, 3, h);
    load_parts(a.w, W, 1, h);
    if (!a.push) load_parts(a.xh, XH, 3, h);
    if (a.push) {
        // particle_tools.f90:58-82: x += u*dt/gamma; x_half = (x_old + x_new)/2
#pragma unroll
        for (int j = 0; j < DEP_G; j++) {
            const double dt_gp = a.dt / sqrt(1.0 + (U[0][j] * U[0][j] + U[1][j] * U[1][j] + U[2][j] * U[2][j]));
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double xn = X[c][j] + U[c][j] * dt_gp;
                XH[c][j] = 0.5 * (X[c][j] + xn);
                X[c][j] = xn;
            }
        }
        if (full) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                st4(a.x + c * ld + ib, dv4{X[c][0], X[c][1], X[c][2], X[c][3]});
                st4(a.xh + c * ld + ib, dv4{XH[c][0], XH[c][1], XH[c][2], XH[c][3]});
            }
        } else {
#pragma unroll
            for (int j = 0; j < DEP_G; j++) {
                const int64_t ip = ib + j;
                if (ip >= a.np) break;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    a.x[c * ld + ip] = X[c][j];
                    a.xh[c * ld + ip] = XH[c][j];
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < DEP_G; j++) {
        if (W[0][j] == 0.0) continue;
        int ix, ir;
        if (doJ && (fabs(U[0][j]) + fabs(U[1][j]) + fabs(U[2][j])) != 0.0 &&
            cell_of(XH[0][j], XH[1][j], XH[2][j], gJ, ix, ir)) {
            bx[0] = min(bx[0], ix); bx[1] = max(bx[1], ix); bx[2] = min(bx[2], ir); bx[3] = max(bx[3], ir);
        }
        if (doR && cell_of(X[0][j], X[1][j], X[2][j], gR, ix, ir)) {
            bx[4] = min(bx[4], ix); bx[5] = max(bx[5], ix); bx[6] = min(bx[6], ir); bx[7] = max(bx[7], ir);
        }
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
    // tiles: one component at a time, (TX x TR nodes) x nmod modes, complex = 2 doubles
    const int TXJ = box[1] >= box[0] ? box[1] - box[0] + 2 : 0, TRJ = box[3] >= box[2] ? box[3] - box[2] + 2 : 0;
    const int TXR = box[5] >= box[4] ? box[5] - box[4] + 2 : 0, TRR = box[7] >= box[6] ? box[7] - box[6] + 2 : 0;
    const int64_t nJ = (int64_t)TXJ * TRJ * nmod * 2, nR = (int64_t)TXR * TRR * nmod * 2;
    const bool tileJ = doJ && nJ > 0 && nJ * (int64_t)sizeof(double) <= DEP_LDS_BYTES;
    const bool tileR = doR && nR > 0 && nR * (int64_t)sizeof(double) <= DEP_LDS_BYTES;
    const bool slotJ = tileJ && TRJ <= DEP_SLOT_R && TXJ <= SW;
    const bool slotR = tileR && TRR <= DEP_SLOT_R && TXR <= SW;

    // ---- phases 2 and 3, one pass per component (J_x, J_y, J_z, rho): zero
    // the tile, deposit into it, then write it to the block's slot (a tile no
    // taller than DEP_SLOT_R rows and no wider than SW columns) or add
    // it to the grid with global atomics.  A thread's consecutive particles
    // usually share a cell (the lattice of a cell is ordered with the
    // azimuthal index innermost), so each thread sums its per-corner
    // contributions in registers while the cell stays the same and adds them
    // to the LDS tile once per cell.
    const int64_t sr = gJ.nx + 1, sm = sr * (gJ.nr + 1), sl = sm * nmod;
    for (int l = 0; l < 4; l++) {
        const bool isJ = l < 3;
        if ((isJ && !doJ) || (!isJ && !doR)) continue;
        const GridGeom& gg = isJ ? gJ : gR;
        double* tile_l = tile;
        C* grid_l = isJ ? a.J + l * sl : a.R;
        const int bx0 = isJ ? box[0] : box[4], br0 = isJ ? box[2] : box[6];
        const int TX = isJ ? TXJ : TXR, TR = isJ ? TRJ : TRR;
        const bool tiled = isJ ? tileJ : tileR;
        const int64_t nt = isJ ? nJ : nR;
        if (tiled) {
            for (int64_t i = tid; i < nt; i += DEP_THREADS) tile[i] = 0.0;
            __syncthreads();
        }
        double2 acc[NMOD][4];   // per-cell sums in FP64 in both precisions
        int cx = -1, cr = 0;
        // a cell's sums go to the tile, or, for a block without one (its box
        // too large), to the grid with one atomic per node value and cell
        auto flush = [&]() __attribute__((always_inline)) {
            if (cx < 0) return;
            if (!tiled) {
#pragma unroll
                for (int m = 0; m < NMOD; m++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        R* dst = reinterpret_cast<R*>(grid_l + (cx + (q & 1)) + (int64_t)(cr + (q >> 1)) * sr + m * sm);
                        unsafeAtomicAdd(dst, (R)acc[m][q].x);
                        unsafeAtomicAdd(dst + 1, (R)acc[m][q].y);
                    }
                return;
            }
            const int tx = cx - bx0, tr = cr - br0;
#pragma unroll
            for (int m = 0; m < NMOD; m++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int t = ((m * TR + (tr + (q >> 1))) * TX + (tx + (q & 1))) * 2;
                    lds_add(tile_l + t, (double)acc[m][q].x);
                    lds_add(tile_l + t + 1, (double)acc[m][q].y);
                }
        };
        for (int h = 0; h < PPT / DEP_G; h++) {
        double P[3][DEP_G], V[3][DEP_G], Wt[3][DEP_G];   // Wt: row 0 only
        load_parts(isJ ? a.xh : a.x, P, 3, h);
        if (isJ) load_parts(a.u, V, 3, h);
        load_parts(a.w, Wt, 1, h);
#pragma unroll
        for (int j = 0; j < DEP_G; j++) {
            const double wp = Wt[0][j];
            if (wp == 0.0) continue;
            double q0;   // weight of this component: w*u_l/gamma for J, w for rho
            PartGeom pg;
            if (isJ) {
                const double u0 = V[0][j], u1 = V[1][j], u2 = V[2][j];
                if ((fabs(u0) + fabs(u1) + fabs(u2)) == 0.0) continue;
                if (!part_geom(pg, P[0][j], P[1][j], P[2][j], gg, -1.0)) continue;
                const double gp = sqrt(1 + u0 * u0 + u1 * u1 + u2 * u2);
                q0 = (l == 0 ? u0 : l == 1 ? u1 : u2) * wp / gp;
            } else {
                if (!part_geom(pg, P[0][j], P[1][j], P[2][j], gg, -1.0)) continue;
                q0 = wp;
            }
            // in-tile test on the cell found here (phase 1 may round differently at a cell edge)
            const int64_t tx = pg.ix - bx0, tr = pg.ir - br0;
            if (!tiled || (tx >= 0 && tx + 1 < TX && tr >= 0 && tr + 1 < TR)) {
                if (pg.ix != cx || pg.ir != cr) {
                    flush();
#pragma unroll
                    for (int m = 0; m < NMOD; m++)
#pragma unroll
                        for (int q = 0; q < 4; q++) acc[m][q] = cmk(0.0, 0.0);
                    cx = (int)pg.ix;
                    cr = (int)pg.ir;
                }
#pragma unroll
                for (int m = 0; m < NMOD; m++) {
                    const double2 pv = crs(q0, pg.ph[m]);
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        acc[m][q] = cadd(acc[m][q], crs(pg.S[0][q & 1] * pg.S[1][q >> 1], pv));
                }
            } else {
#pragma unroll
                for (int m = 0; m < NMOD; m++) {
                    const double2 pv = crs(q0, pg.ph[m]);
                    for (int q = 0; q < 4; q++) {
                        const double2 d = crs(pg.S[0][q & 1] * pg.S[1][q >> 1], pv);
                        R* dst = reinterpret_cast<R*>(grid_l + (pg.ix + (q & 1)) + (pg.ir + (q >> 1)) * sr + m * sm);
                        unsafeAtomicAdd(dst, (R)d.x);
                        unsafeAtomicAdd(dst + 1, (R)d.y);
                    }
                }
            }
        }
        }
        flush();
        if (!tiled) continue;
        __syncthreads();
        const bool slot = isJ ? slotJ : slotR;
        const int64_t n = nt / 2;
        if (slot) {
            // tile index (m*TR + rr)*TX + xx -> slot index ((l*nmod + m)*DEP_SLOT_R + rr)*SW + xx
            // (rho: l = 0 of its own slot array)
            C* sl_base = isJ ? a.slotJ + (int64_t)blk * (3 * nmod * DEP_SLOT_R * SW) +
                                   (int64_t)l * nmod * DEP_SLOT_R * SW
                             : a.slotR + (int64_t)blk * (nmod * DEP_SLOT_R * SW);
            for (int64_t i = tid; i < n; i += DEP_THREADS) {
                const int64_t rest = i / TX;
                sl_base[((rest / TR) * DEP_SLOT_R + rest % TR) * SW + i % TX] =
                    cmake<R>((R)tile[2 * i], (R)tile[2 * i + 1]);
            }
        } else {
            for (int64_t i = tid; i < n; i += DEP_THREADS) {
                const R re = (R)tile[2 * i], im = (R)tile[2 * i + 1];
                if (re == 0 && im == 0) continue;
                const int64_t xx = i % TX, rest = i / TX, rr = rest % TR, m = rest / TR;
                R* dst = reinterpret_cast<R*>(grid_l + (bx0 + xx) + (br0 + rr) * sr + m * sm);
                unsafeAtomicAdd(dst, re);
                unsafeAtomicAdd(dst + 1, im);
            }
        }
        __syncthreads();   // the tile is zeroed again by the next pass
    }

    const int4 none = make_int4(-1, 0, 0, 0);
    if (tid == 0) {
        if (doJ) a.desc[blk] = slotJ ? make_int4(box[2], box[0], TXJ, TRJ) : none;
        if (doR) a.desc[a.nblocks + blk] = slotR ? make_int4(box[6], box[4], TXR, TRR) : none;
    }
}

// ------------------------------------------------------------------ slot combine
//
// Descriptor d (0 <= d < 2*nblocks; J first, then rho) gets the bucket key
// which*nrb + first node row (nrb = nr + 1 buckets per grid), or 2*nrb
// without a slot.
// A stable radix sort of the keys lists every bucket's slots in block order.
__global__ void k_slot_keys(const int4* __restrict__ desc, int nb, int nrb, uint32_t* __restrict__ keys) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= 2 * nb) return;
    const int row = desc[d].x;
    keys[d] = row >= 0 ? (uint32_t)((d >= nb ? nrb : 0) + row) : (uint32_t)(2 * nrb);
}

// bounds[2*k], bounds[2*k+1]: first and one-past-last sorted position of bucket k (zeroed before)
__global__ void k_slot_bounds(const uint32_t* __restrict__ keys, int n, uint32_t nkeys, int32_t* __restrict__ bounds) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t k = keys[p];
    if (k >= nkeys) return;
    if (p == 0 || keys[p - 1] != k) bounds[2 * k] = p;
    if (p == n - 1 || keys[p + 1] != k) bounds[2 * k + 1] = p + 1;
}

constexpr int CMB_T = 256, CMB_LIST = 1024;

// One workgroup per (256 node columns, node row r, grid): the slots whose
// first node row is r - layer (layer < DEP_SLOT_R, their node row `layer` is
// r) and whose columns overlap the workgroup's are listed in LDS in sorted
// order (stable compaction with ballots), with their column range; each
// thread then sums the listed slots that cover its node and adds the sum to
// the grid once.  The order of the sum is fixed: layer 0 first, blocks in
// index order within a layer.  The candidates of a layer come from the bucket
// of row r - layer (sidx and bounds from the sort of the slot keys) or, when
// bounds is null (small launches, where the sort's dozen launches cost more
// than the scan), from all nb descriptors in block order: the same slots in
// the same order, so the two forms give the same sums.
template <int NMOD, typename RT, int SW>
__global__ void __launch_bounds__(CMB_T) k_dep_combine(cplx_t<RT>* __restrict__ J, cplx_t<RT>* __restrict__ R,
                                                        const cplx_t<RT>* __restrict__ slotJ,
                                                        const cplx_t<RT>* __restrict__ slotR,
                                                        const int4* __restrict__ desc, const int32_t* __restrict__ sidx,
                                                        const int32_t* __restrict__ bounds, int nb, int nrb,
                                                        int64_t sr, int64_t sm, int nx1) {
    constexpr int NC = 3 * NMOD;
    const int which = blockIdx.z;
    typedef cplx_t<RT> C;
    C* grid = which ? R : J;
    if (grid == nullptr) return;
    const C* slot = which ? slotR : slotJ;
    const int ncomp = which ? NMOD : NC;
    const int64_t sstride = (int64_t)ncomp * DEP_SLOT_R * SW;
    const int r = blockIdx.y, xs = blockIdx.x * CMB_T, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int x = xs + tid;
    __shared__ int4 list[CMB_LIST];   // (4*block + layer, first column, width, 0)
    __shared__ int wcount[CMB_T / 64];
    __shared__ int nlist;
    C acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = cmake<RT>(0, 0);
    bool hit = false;
    auto drain = [&](int n) __attribute__((always_inline)) {
        for (int e = 0; e < n; e++) {
            const int4 ent = list[e];
            if (x >= ent.y && x < ent.y + ent.z) {
                hit = true;
                const int b = ent.x >> 2, layer = ent.x & 3;
                const C* sp = slot + (int64_t)b * sstride + layer * SW + (x - ent.y);
#pragma unroll
                for (int c = 0; c < NC; c++)
                    if (c < ncomp) acc[c] = cadd(acc[c], sp[c * DEP_SLOT_R * SW]);
            }
        }
    };
    if (tid == 0) nlist = 0;
    __syncthreads();
    for (int layer = 0; layer < DEP_SLOT_R; layer++) {
        const int row = r - layer;
        if (row < 0 || row >= nrb) continue;
        const int k = which * nrb + row;
        const int b0 = bounds ? bounds[2 * k] : 0, b1 = bounds ? bounds[2 * k + 1] : nb;
        for (int q0 = b0; q0 < b1; q0 += CMB_T) {
            const int q = q0 + tid;
            int4 ent = make_int4(-1, 0, 0, 0);
            if (q < b1) {
                const int b = bounds ? sidx[q] - which * nb : q;
                const int4 d = desc[which * nb + b];
                if (d.x == row && layer < d.w && d.y < xs + CMB_T && d.y + d.z > xs)
                    ent = make_int4(4 * b + layer, d.y, d.z, 0);
            }
            // stable compaction: wave ballot + per-wave offsets
            const unsigned long long m = __ballot(ent.x >= 0);
            const int before = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0) wcount[wv] = __popcll(m);
            __syncthreads();
            int off = nlist;
            for (int w = 0; w < wv; w++) off += wcount[w];
            if (ent.x >= 0) list[off + before] = ent;
            __syncthreads();
            if (tid == 0) {
                int tot = 0;
                for (int w = 0; w < CMB_T / 64; w++) tot += wcount[w];
                nlist += tot;
            }
            __syncthreads();
            if (nlist > CMB_LIST - CMB_T) {
                if (x < nx1) drain(nlist);
                __syncthreads();
                if (tid == 0) nlist = 0;
                __syncthreads();
            }
        }
    }
    if (x < nx1) drain(nlist);
    if (x >= nx1 || !hit) return;
    C* gp = grid + x + (int64_t)r * sr;
#pragma unroll
    for (int c = 0; c < NC; c++)
        if (c < ncomp) gp[c * sm] = cadd(gp[c * sm], acc[c]);
}

// Stable compaction: flags, exclusive scan, scatter of every particle array.
__global__ void k_keep_flags(const double* __restrict__ coord, int64_t ld, double l0, double l1, double l2, double l3,
                             int32_t* __restrict__ flag, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    const double x = coord[ip], y = coord[ld + ip], z = coord[2 * ld + ip];
    // rounded as written (no contraction), as in the selection of k_cell_keys
    const double r2 = __dadd_rn(__dmul_rn(y, y), __dmul_rn(z, z));
    flag[ip] = (x >= l0 && x <= l1 && r2 >= l2 && r2 <= l3) ? 1 : 0;
}

template <typename T>
__global__ void k_compact_scatter(const int32_t* __restrict__ flag, const int64_t* __restrict__ pos,
                                  const T* __restrict__ src, T* __restrict__ dst, int64_t ld_src,
                                  int64_t ld_dst, int ncomp, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np || !flag[ip]) return;
    const int64_t o = pos[ip];
    for (int c = 0; c < ncomp; c++) dst[c * ld_dst + o] = src[c * ld_src + ip];
}

// Launch kernel template K<NMOD> for NMOD = nkO + 1 (check_resident bounds nkO).
#define CH_MODE_SWITCH(nkO, LAUNCH)                  \
    switch ((nkO) + 1) {                             \
        case 1: LAUNCH(1); break;                    \
        case 2: LAUNCH(2); break;                    \
        case 3: LAUNCH(3); break;                    \
        case 4: LAUNCH(4); break;                    \
        case 5: LAUNCH(5); break;                    \
        case 6: LAUNCH(6); break;                    \
        case 7: LAUNCH(7); break;                    \
        default: LAUNCH(8); break;                   \
    }

static int32_t check_resident(int np, int64_t ld, int nkO) {
    if (np < 0 || ld < np) return fail(ST_EINVAL, "resident particles: np=%d ld=%lld", np, (long long)ld);
    if (nkO < 0 || nkO >= MAX_MODES)
        return fail(ST_EUNSUP, "azimuthal modes 0..%d requested; this build supports up to %d modes", nkO, MAX_MODES);
    return ST_OK;
}

// get_dens_on_grid (species.py:400-427): dens(nx+1, nr+1, nkO+1) = the
// deposited density (grid_deps.f90:89-147, ghost fold included) times
// VGrid[ir] per radial row; dens is overwritten.
template <typename W>
static int32_t dens_on_grid_impl(const double* coord, const W* wghts, int64_t ld, int32_t np, double* dens,
                                 double leftX, const double* Rgrid, double dx_inv, double dr_inv, double r0,
                                 double rmax, int32_t nx, int32_t nr, int32_t nkO, const double* vgrid) {
    CH_TRY(check_resident(np, ld, nkO));
    if (nx < 1 || nr < 1 || !dens || !vgrid || !is_device_ptr(dens) || !is_device_ptr(vgrid))
        return fail(ST_EINVAL, "dens_on_grid: nx=%d nr=%d; dens and vgrid are device arrays", nx, nr);
    double2* D = reinterpret_cast<double2*>(dens);
    CH_TRY(fill_bytes(D, 0, (size_t)(nx + 1) * (nr + 1) * (nkO + 1) * sizeof(double2)));
    if (np > 0) {
        const GridGeom g = resident_geom(leftX, Rgrid, dx_inv, dr_inv, r0, rmax, nx, nr, nkO);
        CH_TRY(launch_dep_dens<W>(coord, PStride{1, ld}, wghts, D, g, 0, np));
    }
    CH_TRY(ghost_fold<double>(D, nx + 1, nr + 1, nkO + 1));
    return scale_rows(D, vgrid, nx + 1, nr + 1, nkO + 1);
}

// ------------------------------------------------------------------ counting sort by cell
//
// key = ir*nx + ix for a particle in cell (ix, ir) of the grid, nx*nr for a
// kept particle outside the grid, nx*nr + 1 for a particle outside lims (it is
// dropped).  The cell comes from the same expressions as grid_deps.f90:46-47,
// evaluated with explicitly rounded operations so that the integer key of a
// particle does not depend on contraction choices of the compiler.
constexpr int KEYS_T = 256, KEYS_MAX_BLOCKS = 2048;
__global__ void __launch_bounds__(KEYS_T) k_cell_keys(const double* __restrict__ x, const double* __restrict__ xh,
                                                       int64_t ld, int np, double leftX, double dx_inv, double dr_inv,
                                                       double r0, int nx, int nr, int use_lims, double l0, double l1,
                                                       double l2, double l3, int test_half,
                                                       uint32_t* __restrict__ keys, int32_t* __restrict__ kept) {
    __shared__ int32_t wkept[KEYS_T / 64];
    const uint32_t ncell = (uint32_t)nx * (uint32_t)nr;
    int32_t nkept = 0;    // this wave's kept particles (the same in every lane)
    // grid-stride: a bounded number of blocks, one counter update per block
    for (int64_t ip = (int64_t)blockIdx.x * KEYS_T + threadIdx.x; ip - threadIdx.x < np;
         ip += (int64_t)gridDim.x * KEYS_T) {
        const bool valid = ip < np;
        uint32_t key = 0;
        if (valid) {
            const double px = x[ip], py = x[ld + ip], pz = x[2 * ld + ip];
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
        nkept += (int32_t)__popcll(__ballot(valid && key <= ncell));
    }
    if ((threadIdx.x & 63) == 0) wkept[threadIdx.x >> 6] = nkept;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t t = 0;
        for (int w = 0; w < KEYS_T / 64; w++) t += wkept[w];
        if (t) atomicAdd(kept, t);
    }
}

template <typename R>
__global__ void k_gather_particles(const int32_t* __restrict__ idx, int n, const double* __restrict__ x,
                                   const double* __restrict__ xh, const R* __restrict__ u,
                                   const R* __restrict__ w, int64_t ld, double* __restrict__ x2,
                                   double* __restrict__ xh2, R* __restrict__ u2, R* __restrict__ w2,
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

template <typename R>
static int32_t sort_cells_impl(const double* coord, const double* coord_half, const R* momenta, const R* wghts,
                               int64_t ld, int32_t np, double* coord2, double* coord_half2, R* momenta2, R* wghts2,
                               int64_t ld2, double leftX, double dx_inv, double dr_inv, double r0, int32_t nx,
                               int32_t nr, const double* lims, int32_t test_half, int32_t* np_out, int32_t* perm) {
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
    const int64_t kb = (np + KEYS_T - 1) / KEYS_T;
    hipLaunchKernelGGL(k_cell_keys, (unsigned)(kb < KEYS_MAX_BLOCKS ? kb : KEYS_MAX_BLOCKS), KEYS_T, 0, stream(), coord,
                       coord_half, ld, np, leftX, dx_inv,
                       dr_inv, r0, nx, nr, lims ? 1 : 0, L[0], L[1], L[2], L[3], test_half, keys, kept);
    CH_LAUNCH_CHECK("k_cell_keys");
    const uint64_t maxkey = (uint64_t)nx * nr + 1;
    int bits = 1;
    while ((1ull << bits) <= maxkey) bits++;
    CH_TRY(radix_sort_keys(keys, idx, ktmp, itmp, np, bits, 12, 13));
    int32_t hk = 0;
    CH_HIP(hipMemcpyAsync(&hk, kept, sizeof hk, hipMemcpyDeviceToHost, stream()));
    CH_HIP(hipStreamSynchronize(stream()));
    hipLaunchKernelGGL(k_gather_particles<R>, blocks_for(hk, 256), 256, 0, stream(), idx, hk, coord, coord_half, momenta,
                       wghts, ld, coord2, coord_half2, momenta2, wghts2, ld2);
    CH_LAUNCH_CHECK("k_gather_particles");
    *np_out = hk;
    return ST_OK;
}

// The device list of a resident gather-push: none, or the one analytic
// undulator of undul_params (host, 4 doubles), or the described devices.
static int32_t undul_devices(ExtDevices& devs, const double* undul_params) {
    memset(&devs, 0, sizeof devs);
    if (!undul_params) return ST_OK;
    devs.n = 1;
    return make_ext_device(devs.d[0], DEV_UNDUL, undul_params, 4, nullptr, 0, 0.0);
}


template <typename R>
static int32_t gather_push_impl(const double* coord, R* momenta, const R* wghts, int64_t ld, int32_t np,
                                const cplx_t<R>* Fld, double leftX, const double* Rgrid, double dx_inv, double dr_inv,
                                double r0, double rmax, int32_t nx, int32_t nr, int32_t nkO, double push_dt,
                                const ExtDevices& devs, const double* frame) {
    CH_TRY(check_resident(np, ld, nkO));
    if (np == 0) return ST_OK;
    const PStride s = {1, ld};
    GridGeom g = resident_geom(leftX, Rgrid, dx_inv, dr_inv, r0, rmax, nx, nr, nkO);
    g.leftX_dev = frame;
    KTimer kt(KF_GATHER_PUSH);
#define GATHER_LAUNCH(N)                                                                                      \
    hipLaunchKernelGGL((k_gather_push<N, R>), blocks_for(np, GP_T), GP_T, 0, stream(), coord, s, momenta, s, wghts, \
                       Fld, g, push_dt, devs, np)
    CH_MODE_SWITCH(nkO, GATHER_LAUNCH)
#undef GATHER_LAUNCH
    CH_LAUNCH_CHECK("k_gather_push");
    return ST_OK;
}

// Stable removal of the particles outside lims.  With destination arrays
// (dst, leading dimension ld_dst) the kept particles are scattered there and
// the sources are left as they were; without, they are scattered to scratch
// memory and copied back in place.
template <typename R>
static int32_t compact_impl(double* coord, double* coord_half, R* momenta, R* wghts, int64_t ld, int32_t np,
                            const double* lims, int32_t test_half, int32_t* np_out, double* d_coord = nullptr,
                            double* d_half = nullptr, R* d_mom = nullptr, R* d_w = nullptr, int64_t ld_dst = 0) {
    CH_TRY(check_resident(np, ld, 0));
    if (lims == nullptr || np_out == nullptr) return fail(ST_EINVAL, "soa_compact: null lims/np_out");
    if (np == 0) {
        *np_out = 0;
        return ST_OK;
    }
    const bool into = d_coord != nullptr;
    if (into && (!d_half || !d_mom || !d_w || ld_dst < np))
        return fail(ST_EINVAL, "soa_compact_into: four destination arrays of leading dimension >= np");
    int32_t* flag = static_cast<int32_t*>(scratch(4, (size_t)np * sizeof(int32_t)));
    int64_t* pos = static_cast<int64_t*>(scratch(5, ((size_t)np + 8) * sizeof(int64_t)));
    double* tmp = into ? nullptr : static_cast<double*>(scratch(7, (size_t)np * 3 * sizeof(double)));
    if (!flag || !pos || (!into && !tmp)) return ST_ENOMEM;
    int64_t* total = pos + np;
    const double* test = test_half ? coord_half : coord;
    KTimer kt(KF_SORT);
    hipLaunchKernelGGL(k_keep_flags, blocks_for(np, 256), 256, 0, stream(), test, ld, lims[0], lims[1], lims[2],
                       lims[3], flag, np);
    CH_LAUNCH_CHECK("k_keep_flags");
    CH_TRY(exclusive_scan(flag, pos, np, total, 6));
    // each array: scatter the kept particles to the destination, or to the
    // work buffer and back
    auto one = [&](auto* arr, decltype(arr) dst, int nc) -> int32_t {
        typedef typename std::remove_pointer<decltype(arr)>::type T;
        if (into) {
            hipLaunchKernelGGL((k_compact_scatter<T>), blocks_for(np, 256), 256, 0, stream(), flag, pos, arr, dst, ld,
                               ld_dst, nc, np);
            CH_LAUNCH_CHECK("k_compact_scatter");
            return ST_OK;
        }
        T* t = reinterpret_cast<T*>(tmp);
        hipLaunchKernelGGL((k_compact_scatter<T>), blocks_for(np, 256), 256, 0, stream(), flag, pos, arr, t, ld,
                           (int64_t)np, nc, np);
        CH_LAUNCH_CHECK("k_compact_scatter");
        CH_HIP(hipMemcpy2DAsync(arr, ld * sizeof(T), t, (size_t)np * sizeof(T), (size_t)np * sizeof(T), nc,
                                hipMemcpyDeviceToDevice, stream()));
        return ST_OK;
    };
    CH_TRY(one(coord, d_coord, 3));
    CH_TRY(one(coord_half, d_half, 3));
    CH_TRY(one(momenta, d_mom, 3));
    CH_TRY(one(wghts, d_w, 1));
    int64_t h = 0;
    CH_HIP(hipMemcpyAsync(&h, total, sizeof h, hipMemcpyDeviceToHost, stream()));
    CH_HIP(hipStreamSynchronize(stream()));
    *np_out = (int32_t)h;
    return ST_OK;
}

template <typename R>
static int32_t push_deposit_impl(double* coord, double* coord_half, const R* momenta, const R* wghts, int64_t ld,
                                 int32_t np, double dt, double leftX_J, double leftX_rho, const double* Rgrid,
                                 double dx_inv, double dr_inv, double r0, double rmax, int32_t nx, int32_t nr,
                                 int32_t nkO, cplx_t<R>* curr, cplx_t<R>* dens, const double* frame) {
    typedef cplx_t<R> C;
    CH_TRY(check_resident(np, ld, nkO));
    if (np == 0) return ST_OK;
    DepArgsT<R> a;
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
    a.gJ.leftX_dev = frame;
    a.gR.leftX_dev = frame ? frame + 1 : nullptr;
    a.J = curr;
    a.R = dens;
    {
        auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 31) == 0; };
        a.vec4 = (ld % 4 == 0) && al(coord) && al(coord_half) && al(momenta) && al(wghts);
    }
    // 512-particle blocks where the particles average 8 or more per cell
    const bool big = (int64_t)np >= (int64_t)8 * nx * nr;
    const int ppt = big ? 8 : 4, sw = 8 * ppt;
    const int nb = (int)blocks_for(np, DEP_THREADS * ppt);
    const int nmod = nkO + 1, nrb = nr + 1;
    a.nblocks = nb;
    a.slotJ = a.J ? static_cast<C*>(scratch(16, (size_t)nb * 3 * nmod * DEP_SLOT_R * sw * sizeof(C))) : nullptr;
    a.slotR = a.R ? static_cast<C*>(scratch(17, (size_t)nb * nmod * DEP_SLOT_R * sw * sizeof(C))) : nullptr;
    a.desc = static_cast<int4*>(scratch(18, (size_t)2 * nb * sizeof(int4)));
    uint32_t* keys = static_cast<uint32_t*>(scratch(19, (size_t)2 * nb * 4 * sizeof(uint32_t)));
    int32_t* bounds = static_cast<int32_t*>(scratch(23, (size_t)2 * 2 * nrb * sizeof(int32_t)));
    if ((a.J && !a.slotJ) || (a.R && !a.slotR) || !a.desc || !keys || !bounds) return ST_ENOMEM;
    if (!a.J && !a.R) return ST_OK;
    KTimer kt(a.J ? KF_DEP_CURR : KF_DEP_DENS);
    const int64_t sr = nx + 1, sm = sr * (nr + 1);
    dim3 grid((unsigned)((nx + 1 + CMB_T - 1) / CMB_T), (unsigned)(nr + 1), 2u);
    // small launches: every combine workgroup scans all descriptors
    // (DEP_SLOT_R x nb / CMB_T rounds of coalesced L2 reads, 64 B per
    // descriptor and workgroup) instead of sorting them into buckets, which
    // takes a dozen dependent launches of about 5 us each; up to 512 blocks
    // (131072 particles), or while the scans read at most 512 MiB in all (L2
    // reads: on a 2048 x 256 bunch of 300 k particles, 352 MiB of them take
    // less time than the sort)
    const bool scan_all = nb <= 512 || (int64_t)nb * grid.x * grid.y * grid.z <= ((int64_t)1 << 23);
#define CMB_LAUNCH(N, IDX, BOUNDS)                                                                                \
    do {                                                                                                          \
        if (big)                                                                                                  \
            hipLaunchKernelGGL((k_dep_combine<N, R, 64>), grid, CMB_T, 0, stream(), a.J, a.R, a.slotJ, a.slotR,   \
                               a.desc, IDX, BOUNDS, nb, nrb, sr, sm, nx + 1);                                     \
        else                                                                                                      \
            hipLaunchKernelGGL((k_dep_combine<N, R, 32>), grid, CMB_T, 0, stream(), a.J, a.R, a.slotJ, a.slotR,   \
                               a.desc, IDX, BOUNDS, nb, nrb, sr, sm, nx + 1);                                     \
    } while (0)
    if (scan_all) {
        // every descriptor of a deposited grid is written by its block
#define DEP_LAUNCH(N)                                                                                             \
    do {                                                                                                          \
        if (big)                                                                                                  \
            hipLaunchKernelGGL((k_push_deposit<N, R, 8>), nb, DEP_THREADS, 0, stream(), a);                       \
        else                                                                                                      \
            hipLaunchKernelGGL((k_push_deposit<N, R, 4>), nb, DEP_THREADS, 0, stream(), a);                       \
    } while (0)
        CH_MODE_SWITCH(nkO, DEP_LAUNCH)
        CH_LAUNCH_CHECK("k_push_deposit");
#define CMB_SCAN(N) CMB_LAUNCH(N, nullptr, nullptr)
        CH_MODE_SWITCH(nkO, CMB_SCAN)
#undef CMB_SCAN
        CH_LAUNCH_CHECK("k_dep_combine");
        return ST_OK;
    }
    // descriptors of a grid not deposited here read "no slot" (row -1)
    CH_TRY(fill_bytes(a.desc, 0xFF, (size_t)2 * nb * sizeof(int4)));
    CH_MODE_SWITCH(nkO, DEP_LAUNCH)
#undef DEP_LAUNCH
    CH_LAUNCH_CHECK("k_push_deposit");
    // bucket the slots by (grid, cell row), in block order
    const int nd = 2 * nb;
    uint32_t* keys_tmp = keys + nd;
    int32_t* idx = reinterpret_cast<int32_t*>(keys + 2 * (size_t)nd);
    int32_t* idx_tmp = idx + nd;
    hipLaunchKernelGGL(k_slot_keys, blocks_for(nd, 256), 256, 0, stream(), a.desc, nb, nrb, keys);
    CH_LAUNCH_CHECK("k_slot_keys");
    int bits = 1;
    while ((1u << bits) <= (uint32_t)(2 * nrb)) bits++;
    CH_TRY(radix_sort_keys(keys, idx, keys_tmp, idx_tmp, nd, bits, 20, 21));
    CH_TRY(fill_bytes(bounds, 0, (size_t)2 * 2 * nrb * sizeof(int32_t)));
    hipLaunchKernelGGL(k_slot_bounds, blocks_for(nd, 256), 256, 0, stream(), keys, nd, (uint32_t)(2 * nrb), bounds);
    CH_LAUNCH_CHECK("k_slot_bounds");
#define CMB_SORTED(N) CMB_LAUNCH(N, idx, bounds)
    CH_MODE_SWITCH(nkO, CMB_SORTED)
#undef CMB_SORTED
#undef CMB_LAUNCH
    CH_LAUNCH_CHECK("k_dep_combine");
    return ST_OK;
}

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

int32_t chimera_soa_dens_on_grid(const double* coord, const double* wghts, int64_t ld, int32_t np, double* dens,
                                 double leftX, const double* Rgrid, double dx_inv, double dr_inv, double r0,
                                 double rmax, int32_t nx, int32_t nr, int32_t nkO, const double* vgrid) {
    return dens_on_grid_impl<double>(coord, wghts, ld, np, dens, leftX, Rgrid, dx_inv, dr_inv, r0, rmax, nx, nr, nkO,
                                     vgrid);
}

int32_t chimera_soa_dens_on_grid_mixed(const double* coord, const float* wghts, int64_t ld, int32_t np,
                                       double* dens, double leftX, const double* Rgrid, double dx_inv, double dr_inv,
                                       double r0, double rmax, int32_t nx, int32_t nr, int32_t nkO,
                                       const double* vgrid) {
    return dens_on_grid_impl<float>(coord, wghts, ld, np, dens, leftX, Rgrid, dx_inv, dr_inv, r0, rmax, nx, nr, nkO,
                                    vgrid);
}

int32_t chimera_soa_gather_push(const double* coord, double* momenta, const double* wghts, int64_t ld, int32_t np,
                                const double* Fld, double leftX, const double* Rgrid, double dx_inv, double dr_inv,
                                double r0, double rmax, int32_t nx, int32_t nr, int32_t nkO, double push_dt,
                                const double* undul_params, const double* frame) {
    ExtDevices devs;
    CH_TRY(undul_devices(devs, undul_params));
    return gather_push_impl<double>(coord, momenta, wghts, ld, np, reinterpret_cast<const double2*>(Fld), leftX, Rgrid,
                                    dx_inv, dr_inv, r0, rmax, nx, nr, nkO, push_dt, devs, frame);
}

int32_t chimera_soa_gather_push_mixed(const double* coord, float* momenta, const float* wghts, int64_t ld, int32_t np,
                                      const float* Fld, double leftX, const double* Rgrid, double dx_inv,
                                      double dr_inv, double r0, double rmax, int32_t nx, int32_t nr, int32_t nkO,
                                      double push_dt, const double* undul_params, const double* frame) {
    ExtDevices devs;
    CH_TRY(undul_devices(devs, undul_params));
    return gather_push_impl<float>(coord, momenta, wghts, ld, np, reinterpret_cast<const float2*>(Fld), leftX, Rgrid,
                                   dx_inv, dr_inv, r0, rmax, nx, nr, nkO, push_dt, devs, frame);
}

int32_t chimera_soa_gather_push_dev(const double* coord, double* momenta, const double* wghts, int64_t ld,
                                    int32_t np, const double* Fld, double leftX, const double* Rgrid, double dx_inv,
                                    double dr_inv, double r0, double rmax, int32_t nx, int32_t nr, int32_t nkO,
                                    double push_dt, const chimera_device_t* devices, int32_t ndev, double t,
                                    const double* frame) {
    ExtDevices devs;
    CH_TRY(ext_devices(devs, devices, ndev, t));
    return gather_push_impl<double>(coord, momenta, wghts, ld, np, reinterpret_cast<const double2*>(Fld), leftX, Rgrid,
                                    dx_inv, dr_inv, r0, rmax, nx, nr, nkO, push_dt, devs, frame);
}

int32_t chimera_soa_gather_push_dev_mixed(const double* coord, float* momenta, const float* wghts, int64_t ld,
                                          int32_t np, const float* Fld, double leftX, const double* Rgrid,
                                          double dx_inv, double dr_inv, double r0, double rmax, int32_t nx, int32_t nr,
                                          int32_t nkO, double push_dt, const chimera_device_t* devices, int32_t ndev,
                                          double t, const double* frame) {
    ExtDevices devs;
    CH_TRY(ext_devices(devs, devices, ndev, t));
    return gather_push_impl<float>(coord, momenta, wghts, ld, np, reinterpret_cast<const float2*>(Fld), leftX, Rgrid,
                                   dx_inv, dr_inv, r0, rmax, nx, nr, nkO, push_dt, devs, frame);
}

// Advance of a device-side window frame, as move_frame forms the new edge
// (chimera_main.py:286-291: Xgrid + shiftX, leftX = Xgrid[0]).
__global__ void k_frame_step(double* __restrict__ f, double shift1, double shift2) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double a = f[1] + shift1;
    f[0] = a;
    f[1] = a + shift2;
}

int32_t chimera_frame_step(double* frame, double shift1, double shift2) {
    if (!is_device_ptr(frame)) return fail(ST_EINVAL, "frame_step: the frame is a device pointer");
    hipLaunchKernelGGL(k_frame_step, 1, 64, 0, stream(), frame, shift1, shift2);
    CH_LAUNCH_CHECK("k_frame_step");
    return ST_OK;
}

int32_t chimera_soa_compact(double* coord, double* coord_half, double* momenta, double* wghts, int64_t ld, int32_t np,
                            const double* lims, int32_t test_half, int32_t* np_out) {
    return compact_impl<double>(coord, coord_half, momenta, wghts, ld, np, lims, test_half, np_out);
}

int32_t chimera_soa_compact_mixed(double* coord, double* coord_half, float* momenta, float* wghts, int64_t ld,
                                  int32_t np, const double* lims, int32_t test_half, int32_t* np_out) {
    return compact_impl<float>(coord, coord_half, momenta, wghts, ld, np, lims, test_half, np_out);
}

int32_t chimera_soa_compact_into(const double* coord, const double* coord_half, const double* momenta,
                                 const double* wghts, int64_t ld, int32_t np, double* d_coord, double* d_half,
                                 double* d_mom, double* d_w, int64_t ld_dst, const double* lims, int32_t test_half,
                                 int32_t* np_out) {
    if (!d_coord) return fail(ST_EINVAL, "soa_compact_into: null destination");
    return compact_impl<double>(const_cast<double*>(coord), const_cast<double*>(coord_half),
                                const_cast<double*>(momenta), const_cast<double*>(wghts), ld, np, lims, test_half,
                                np_out, d_coord, d_half, d_mom, d_w, ld_dst);
}

int32_t chimera_soa_compact_into_mixed(const double* coord, const double* coord_half, const float* momenta,
                                       const float* wghts, int64_t ld, int32_t np, double* d_coord, double* d_half,
                                       float* d_mom, float* d_w, int64_t ld_dst, const double* lims,
                                       int32_t test_half, int32_t* np_out) {
    if (!d_coord) return fail(ST_EINVAL, "soa_compact_into: null destination");
    return compact_impl<float>(const_cast<double*>(coord), const_cast<double*>(coord_half),
                               const_cast<float*>(momenta), const_cast<float*>(wghts), ld, np, lims, test_half,
                               np_out, d_coord, d_half, d_mom, d_w, ld_dst);
}


int32_t chimera_soa_push_deposit(double* coord, double* coord_half, const double* momenta, const double* wghts,
                                 int64_t ld, int32_t np, double dt, double leftX_J, double leftX_rho,
                                 const double* Rgrid, double dx_inv, double dr_inv, double r0, double rmax, int32_t nx,
                                 int32_t nr, int32_t nkO, double* curr, double* dens, const double* frame) {
    return push_deposit_impl<double>(coord, coord_half, momenta, wghts, ld, np, dt, leftX_J, leftX_rho, Rgrid, dx_inv,
                                     dr_inv, r0, rmax, nx, nr, nkO, reinterpret_cast<double2*>(curr),
                                     reinterpret_cast<double2*>(dens), frame);
}

int32_t chimera_soa_push_deposit_mixed(double* coord, double* coord_half, const float* momenta, const float* wghts,
                                       int64_t ld, int32_t np, double dt, double leftX_J, double leftX_rho,
                                       const double* Rgrid, double dx_inv, double dr_inv, double r0, double rmax,
                                       int32_t nx, int32_t nr, int32_t nkO, float* curr, float* dens,
                                       const double* frame) {
    return push_deposit_impl<float>(coord, coord_half, momenta, wghts, ld, np, dt, leftX_J, leftX_rho, Rgrid, dx_inv,
                                    dr_inv, r0, rmax, nx, nr, nkO, reinterpret_cast<float2*>(curr),
                                    reinterpret_cast<float2*>(dens), frame);
}


int32_t chimera_soa_sort_cells(const double* coord, const double* coord_half, const double* momenta,
                               const double* wghts, int64_t ld, int32_t np, double* coord2, double* coord_half2,
                               double* momenta2, double* wghts2, int64_t ld2, double leftX, double dx_inv,
                               double dr_inv, double r0, int32_t nx, int32_t nr, const double* lims, int32_t test_half,
                               int32_t* np_out, int32_t* perm) {
    return sort_cells_impl<double>(coord, coord_half, momenta, wghts, ld, np, coord2, coord_half2, momenta2, wghts2,
                                   ld2, leftX, dx_inv, dr_inv, r0, nx, nr, lims, test_half, np_out, perm);
}

int32_t chimera_soa_sort_cells_mixed(const double* coord, const double* coord_half, const float* momenta,
                                     const float* wghts, int64_t ld, int32_t np, double* coord2, double* coord_half2,
                                     float* momenta2, float* wghts2, int64_t ld2, double leftX, double dx_inv,
                                     double dr_inv, double r0, int32_t nx, int32_t nr, const double* lims,
                                     int32_t test_half, int32_t* np_out, int32_t* perm) {
    return sort_cells_impl<float>(coord, coord_half, momenta, wghts, ld, np, coord2, coord_half2, momenta2, wghts2,
                                  ld2, leftX, dx_inv, dr_inv, r0, nx, nr, lims, test_half, np_out, perm);
}

}  // extern "C"
