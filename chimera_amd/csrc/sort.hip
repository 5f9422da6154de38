// sort.hip -- particle bookkeeping of particle_tools.f90 (domain filter,
// chunk binning, permutation gather) and the integer scan they rest on.
#include "common.hpp"
#include "sort.hpp"
#include "../../include/chimera_hip.h"

namespace chimera {

// ------------------------------------------------------------------ exclusive scan (int64 totals)

constexpr int SCAN_T = 256, SCAN_ITEMS = 4, SCAN_TILE = SCAN_T * SCAN_ITEMS;

__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* sh, int64_t* total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        const int64_t a = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    const int64_t incl = sh[t];
    *total = sh[SCAN_T - 1];
    __syncthreads();
    return incl - v;
}

// Pass 1: sum of each tile.
__global__ void k_tile_sums(const int32_t* __restrict__ in, int64_t n, int64_t* __restrict__ sums) {
    __shared__ int64_t sh[SCAN_T];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int64_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; i++)
        if (base + i < n) s += in[base + i];
    int64_t tot;
    block_excl_scan(s, sh, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Single-block exclusive scan of the tile sums (ntiles <= SCAN_TILE*SCAN_TILE handled by looping).
__global__ void k_scan_sums(int64_t* __restrict__ sums, int64_t ntiles, int64_t* __restrict__ total) {
    __shared__ int64_t sh[SCAN_T];
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < ntiles; b0 += SCAN_TILE) {
        const int64_t base = b0 + (int64_t)threadIdx.x * SCAN_ITEMS;
        int64_t v[SCAN_ITEMS], s = 0;
        for (int i = 0; i < SCAN_ITEMS; i++) {
            v[i] = base + i < ntiles ? sums[base + i] : 0;
            s += v[i];
        }
        int64_t tot;
        int64_t ex = block_excl_scan(s, sh, &tot) + carry;
        for (int i = 0; i < SCAN_ITEMS; i++) {
            if (base + i < ntiles) sums[base + i] = ex;
            ex += v[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

// Pass 3: scan inside each tile, offset by the tile's exclusive prefix.
__global__ void k_tile_scan(const int32_t* __restrict__ in, int64_t n, const int64_t* __restrict__ sums,
                            int64_t* __restrict__ out) {
    __shared__ int64_t sh[SCAN_T];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int64_t v[SCAN_ITEMS], s = 0;
    for (int i = 0; i < SCAN_ITEMS; i++) {
        v[i] = base + i < n ? in[base + i] : 0;
        s += v[i];
    }
    int64_t tot;
    int64_t ex = block_excl_scan(s, sh, &tot) + sums[blockIdx.x];
    for (int i = 0; i < SCAN_ITEMS; i++) {
        if (base + i < n) out[base + i] = ex;
        ex += v[i];
    }
}

int32_t exclusive_scan(const int32_t* in, int64_t* out, int64_t n, int64_t* total_dev, int scratch_slot) {
    if (n <= 0) return ST_OK;
    const int64_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    int64_t* sums = static_cast<int64_t*>(scratch(scratch_slot, (size_t)ntiles * sizeof(int64_t)));
    if (!sums) return ST_ENOMEM;
    hipLaunchKernelGGL(k_tile_sums, (unsigned)ntiles, SCAN_T, 0, stream(), in, n, sums);
    CH_LAUNCH_CHECK("k_tile_sums");
    hipLaunchKernelGGL(k_scan_sums, 1, SCAN_T, 0, stream(), sums, ntiles, total_dev);
    CH_LAUNCH_CHECK("k_scan_sums");
    hipLaunchKernelGGL(k_tile_scan, (unsigned)ntiles, SCAN_T, 0, stream(), in, n, sums, out);
    CH_LAUNCH_CHECK("k_tile_scan");
    return ST_OK;
}

// ------------------------------------------------------------------ sortpartsout (particle_tools.f90:130-153)

__device__ __forceinline__ bool in_domain(const double* c, PStride s, int64_t ip, const double* lims) {
    const double x = c[s.at(ip, 0)], y = c[s.at(ip, 1)], z = c[s.at(ip, 2)];
    const double r2 = y * y + z * z;
    return x >= lims[0] && x <= lims[1] && r2 >= lims[2] && r2 <= lims[3];
}

__global__ void k_stay_flags(const double* __restrict__ coord, PStride s, const double* __restrict__ lims,
                             int32_t* __restrict__ flag, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    flag[ip] = in_domain(coord, s, ip, lims) ? 1 : 0;
}

__global__ void k_stay_scatter(const int32_t* __restrict__ flag, const int64_t* __restrict__ pos,
                               int32_t* __restrict__ idx, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    if (flag[ip]) idx[pos[ip]] = ip;
}

__global__ void k_fill_i32(int32_t* __restrict__ a, int32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ------------------------------------------------------------------ chunk_coords_boundaries (:155-208)

__global__ void k_chunk_ids(const double* __restrict__ coord, PStride s, const double* __restrict__ lims, double X0,
                            double inv, int nchnk, int8_t* __restrict__ ids, int32_t* __restrict__ counts,
                            int32_t* __restrict__ goout, int np) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    const double x = coord[s.at(ip, 0)];
    int8_t id = -2;
    if (in_domain(coord, s, ip, lims)) {
        const double f = floor((x - X0) * inv);
        const int ic = f >= -128.0 && f < 128.0 ? (int)f : -128;
        if (ic >= 0 && ic < nchnk) atomicAdd(&counts[ic], 1);
        id = (int8_t)ic;
    } else {
        atomicAdd(goout, 1);
    }
    ids[ip] = id;
}

// ------------------------------------------------------------------ align_data (:270-324)

__global__ void k_gather_rows(const double* __restrict__ dat, const int64_t* __restrict__ idx, double* __restrict__ tmp,
                              int ncomp, int n, int64_t np0, int32_t* __restrict__ bad) {
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const int64_t j = idx[ip];
    if (j < 0 || j >= np0) {
        atomicOr(bad, 1);
        for (int c = 0; c < ncomp; c++) tmp[(int64_t)ncomp * ip + c] = 0.0;
        return;
    }
    for (int c = 0; c < ncomp; c++) tmp[(int64_t)ncomp * ip + c] = dat[ncomp * j + c];
}

}  // namespace chimera

using namespace chimera;

extern "C" {

int32_t chimera_sortpartsout(int32_t* indx2stay, int32_t* num2stay, const double* coord, const double* lims,
                             int32_t np) {
    if (np < 0 || num2stay == nullptr) return fail(ST_EINVAL, "sortpartsout: bad arguments");
    if (np == 0) {
        *num2stay = 0;
        return ST_OK;
    }
    Stage st;
    int32_t* idx = st.out(indx2stay, (size_t)np);
    const double* c = st.in(coord, 3 * (size_t)np);
    const double* L = st.in(lims, 4);
    CH_TRY(st.status());
    int32_t* flag = static_cast<int32_t*>(scratch(4, (size_t)np * sizeof(int32_t)));
    int64_t* pos = static_cast<int64_t*>(scratch(5, (size_t)np * sizeof(int64_t) + 64));
    if (!flag || !pos) return ST_ENOMEM;
    int64_t* total = pos + np;
    const PStride s = {3, 1};
    hipLaunchKernelGGL(k_stay_flags, blocks_for(np, 256), 256, 0, stream(), c, s, L, flag, np);
    CH_LAUNCH_CHECK("k_stay_flags");
    CH_TRY(exclusive_scan(flag, pos, np, total, 6));
    hipLaunchKernelGGL(k_fill_i32, blocks_for(np, 256), 256, 0, stream(), idx, 0, (int64_t)np);
    CH_LAUNCH_CHECK("k_fill_i32");
    hipLaunchKernelGGL(k_stay_scatter, blocks_for(np, 256), 256, 0, stream(), flag, pos, idx, np);
    CH_LAUNCH_CHECK("k_stay_scatter");
    int64_t h = 0;
    CH_HIP(hipMemcpyAsync(&h, total, sizeof h, hipMemcpyDeviceToHost, stream()));
    CH_HIP(hipStreamSynchronize(stream()));
    *num2stay = (int32_t)h;
    return st.finish();
}

int32_t chimera_chunk_coords_boundaries(int8_t* chunked_indx, int32_t* IndInChnk, int32_t* GoOut, const double* coord,
                                        const double* lims, const double* Xgrid, int32_t nchnk, int32_t np,
                                        int32_t nx) {
    if (np < 0 || nchnk < 1 || nx < 1 || IndInChnk == nullptr || GoOut == nullptr)
        return fail(ST_EINVAL, "chunk_coords_boundaries: bad arguments");
    double X[2];
    const int iend = nchnk > 1 ? (nx + 1) / nchnk : nx;
    if (is_device_ptr(Xgrid)) {
        CH_HIP(hipMemcpy(&X[0], Xgrid, sizeof(double), hipMemcpyDeviceToHost));
        CH_HIP(hipMemcpy(&X[1], Xgrid + iend, sizeof(double), hipMemcpyDeviceToHost));
    } else {
        X[0] = Xgrid[0];
        X[1] = Xgrid[iend];
    }
    const double inv = 1. / (X[1] - X[0]);
    Stage st;
    int8_t* ids = st.out(chunked_indx, (size_t)np);
    const double* c = st.in(coord, 3 * (size_t)np);
    const double* L = st.in(lims, 4);
    CH_TRY(st.status());
    int32_t* counts = static_cast<int32_t*>(scratch(4, ((size_t)nchnk + 1) * sizeof(int32_t)));
    if (!counts) return ST_ENOMEM;
    CH_HIP(hipMemsetAsync(counts, 0, ((size_t)nchnk + 1) * sizeof(int32_t), stream()));
    const PStride s = {3, 1};
    if (np > 0) {
        hipLaunchKernelGGL(k_chunk_ids, blocks_for(np, 256), 256, 0, stream(), c, s, L, X[0], inv, nchnk, ids, counts,
                           counts + nchnk, np);
        CH_LAUNCH_CHECK("k_chunk_ids");
    }
    std::vector<int32_t> h(nchnk + 1);
    CH_HIP(hipMemcpyAsync(h.data(), counts, ((size_t)nchnk + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, stream()));
    CH_HIP(hipStreamSynchronize(stream()));
    IndInChnk[0] = 0;
    for (int i = 0; i < nchnk; i++) IndInChnk[i + 1] = IndInChnk[i] + h[i];
    *GoOut = h[nchnk];
    return st.finish();
}

static int32_t align_entry(double* dat, const int64_t* idx, int32_t np, int32_t np0, int ncomp) {
    if (np < 0 || np0 < 0 || np > np0) return fail(ST_EINVAL, "align_data: np=%d np0=%d", np, np0);
    if (np == 0) return ST_OK;
    Stage st;
    double* d = st.inout(dat, (size_t)ncomp * np0);
    const int64_t* ix = st.in(idx, (size_t)np);
    CH_TRY(st.status());
    double* tmp = static_cast<double*>(scratch(4, (size_t)ncomp * np * sizeof(double) + 64));
    if (!tmp) return ST_ENOMEM;
    int32_t* bad = reinterpret_cast<int32_t*>(tmp + (size_t)ncomp * np);
    CH_HIP(hipMemsetAsync(bad, 0, sizeof(int32_t), stream()));
    hipLaunchKernelGGL(k_gather_rows, blocks_for(np, 256), 256, 0, stream(), d, ix, tmp, ncomp, np, (int64_t)np0, bad);
    CH_LAUNCH_CHECK("k_gather_rows");
    CH_HIP(hipMemcpyAsync(d, tmp, (size_t)ncomp * np * sizeof(double), hipMemcpyDeviceToDevice, stream()));
    int32_t hb = 0;
    CH_HIP(hipMemcpyAsync(&hb, bad, sizeof hb, hipMemcpyDeviceToHost, stream()));
    CH_TRY(st.finish());
    CH_HIP(hipStreamSynchronize(stream()));
    if (hb) return fail(ST_EINVAL, "align_data: an index lies outside 0..%d", np0 - 1);
    return ST_OK;
}

int32_t chimera_align_data_vec(double* dat, const int64_t* chunked_indx, int32_t np, int32_t np0) {
    return align_entry(dat, chunked_indx, np, np0, 3);
}

int32_t chimera_align_data_scl(double* dat, const int64_t* chunked_indx, int32_t np, int32_t np0) {
    return align_entry(dat, chunked_indx, np, np0, 1);
}

}  // extern "C"
