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

// ------------------------------------------------------------------ stable LSD radix sort of integer keys
//
// Keys are sorted RS_BITS bits at a time.  Every pass: (1) each block of
// RS_TILE consecutive items counts its digits, (2) an exclusive scan over the
// (digit, block) counts in digit-major order gives every block its output
// offset per digit, (3) each block writes its items to their offsets in input
// order.  Ranks inside a wave come from a ballot-based match of equal digits
// (the lanes of one digit and their popcount), ranks across the block's waves
// from per-wave digit counts in LDS, so the sort is stable and deterministic.
constexpr int RS_BITS = 8, RS_BINS = 1 << RS_BITS, RS_T = 256, RS_ITEMS = 8, RS_TILE = RS_T * RS_ITEMS;

__device__ __forceinline__ uint64_t match_digit(int d, bool valid) {
    // lanes with the same (valid, digit) as this lane
    uint64_t m = __ballot(1);
    const uint64_t v = __ballot(valid);
    m &= valid ? v : ~v;
#pragma unroll
    for (int b = 0; b < RS_BITS; b++) {
        const bool bit = (d >> b) & 1;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__global__ void __launch_bounds__(RS_T) k_rs_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                   int32_t* __restrict__ hist, int64_t nblocks) {
    __shared__ int32_t h[RS_BINS];
    for (int i = threadIdx.x; i < RS_BINS; i += RS_T) h[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int it = 0; it < RS_ITEMS; it++) {
        const int64_t i = base + (int64_t)it * RS_T + threadIdx.x;
        const bool valid = i < n;
        const int d = valid ? (int)((keys[i] >> shift) & (RS_BINS - 1)) : 0;
        const uint64_t m = match_digit(d, valid);
        const int lead = __ffsll((unsigned long long)m) - 1;
        if (valid && lane_id() == lead) atomicAdd(&h[d], __popcll(m));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < RS_BINS; i += RS_T) hist[(int64_t)i * nblocks + blockIdx.x] = h[i];
}

__global__ void __launch_bounds__(RS_T) k_rs_scatter(const uint32_t* __restrict__ keys, const int32_t* __restrict__ idx,
                                                      int64_t n, int shift, const int64_t* __restrict__ offs,
                                                      int64_t nblocks, uint32_t* __restrict__ keys_out,
                                                      int32_t* __restrict__ idx_out) {
    __shared__ int32_t running[RS_BINS];
    __shared__ int32_t wcount[RS_T / 64][RS_BINS];
    for (int i = threadIdx.x; i < RS_BINS; i += RS_T) running[i] = 0;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    const uint64_t lt = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    for (int it = 0; it < RS_ITEMS; it++) {
        for (int i = threadIdx.x; i < (RS_T / 64) * RS_BINS; i += RS_T) (&wcount[0][0])[i] = 0;
        __syncthreads();
        const int64_t i = base + (int64_t)it * RS_T + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? keys[i] : 0u;
        const int d = (int)((key >> shift) & (RS_BINS - 1));
        const uint64_t m = match_digit(d, valid);
        const int rank = __popcll(m & lt);
        if (valid && rank == 0) wcount[wave][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            int pos = running[d] + rank;
            for (int w = 0; w < wave; w++) pos += wcount[w][d];
            const int64_t o = offs[(int64_t)d * nblocks + blockIdx.x] + pos;
            keys_out[o] = key;
            idx_out[o] = idx ? idx[i] : (int32_t)i;
        }
        __syncthreads();
        for (int b = threadIdx.x; b < RS_BINS; b += RS_T) {
            int add = 0;
            for (int w = 0; w < RS_T / 64; w++) add += wcount[w][b];
            running[b] += add;
        }
        __syncthreads();
    }
}

int32_t radix_sort_keys(uint32_t* keys, int32_t* idx, uint32_t* keys_tmp, int32_t* idx_tmp, int64_t n, int key_bits,
                        int scratch_hist, int scratch_scan) {
    if (n <= 0) return ST_OK;
    const int64_t nblocks = (n + RS_TILE - 1) / RS_TILE;
    int32_t* hist = static_cast<int32_t*>(scratch(scratch_hist, (size_t)RS_BINS * nblocks * sizeof(int32_t)));
    int64_t* offs = static_cast<int64_t*>(scratch(scratch_scan, ((size_t)RS_BINS * nblocks + 8) * sizeof(int64_t)));
    if (!hist || !offs) return ST_ENOMEM;
    uint32_t* kbuf[2] = {keys, keys_tmp};
    int32_t* ibuf[2] = {idx, idx_tmp};
    int passes = (key_bits + RS_BITS - 1) / RS_BITS;
    if (passes < 1) passes = 1;
    for (int p = 0; p < passes; p++) {
        const int src = p & 1, dst = src ^ 1;
        const int shift = p * RS_BITS;
        hipLaunchKernelGGL(k_rs_hist, (unsigned)nblocks, RS_T, 0, stream(), kbuf[src], n, shift, hist, nblocks);
        CH_LAUNCH_CHECK("k_rs_hist");
        CH_TRY(exclusive_scan(hist, offs, (int64_t)RS_BINS * nblocks, nullptr, scratch_scan + 1));
        // pass 0 reads the identity permutation
        hipLaunchKernelGGL(k_rs_scatter, (unsigned)nblocks, RS_T, 0, stream(), kbuf[src], p == 0 ? nullptr : ibuf[src],
                           n, shift, offs, nblocks, kbuf[dst], ibuf[dst]);
        CH_LAUNCH_CHECK("k_rs_scatter");
    }
    if (passes & 1) {  // the result is in the work buffers
        CH_HIP(hipMemcpyAsync(keys, keys_tmp, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream()));
        CH_HIP(hipMemcpyAsync(idx, idx_tmp, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToDevice, stream()));
    }
    return ST_OK;
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
