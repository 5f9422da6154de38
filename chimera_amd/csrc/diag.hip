// diag.hip -- diagnostics on the device: the beam moments behind the envelopes
// of get_beam_envelops (diagnostics.py:177-207), so that a served simulation
// can return its beam envelope without copying the particles to the host, the
// EM energy per kx of nrg_out (diagnostics.py:109-122) and the laser power
// per x of pwr_out (diagnostics.py:124-149).
//
// The 13 weighted sums the envelope formulas need,
//   S = [sum w, sum x w, sum y w, sum z w, sum x^2 w, sum y^2 w, sum z^2 w,
//        sum px^2 w, sum py^2 w, sum pz^2 w, sum x px w, sum y py w, sum z pz w]
// (x, y, z the half-step positions; products formed as the reference writes
// them, x**2*w and x*px*w), are accumulated in FP64 in two launches: a
// grid-stride pass that leaves one partial sum per workgroup, then one
// workgroup that adds the partials.  Both reductions run in a fixed order for
// a given particle count, so the sums are the same bits on every call.
#include <algorithm>
#include <cmath>
#include <complex>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "../../include/chimera_hip.h"

namespace chimera {
constexpr int MOM_T = 256, MOM_NS = 13, MOM_MAXB = 512;
}

// Scratch of chimera_beam_moments_noalloc: the per-workgroup partial sums of
// the first reduction pass, MOM_MAXB rows of MOM_NS doubles (53,248 bytes), in
// every GPU's code object for the life of the process.  Host code never uses
// the name itself (on the host it is only a stand-in for the symbol): the
// device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[chimera::MOM_MAXB * chimera::MOM_NS];

namespace chimera {

// Tree sum of v over the workgroup into sh[0]; every thread must call it.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = MOM_T / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

template <typename U>
__global__ void __launch_bounds__(MOM_T) k_moments_part(const double* __restrict__ xh, const U* __restrict__ u,
                                                        const U* __restrict__ w, int64_t ld, int np,
                                                        double* __restrict__ part) {
    __shared__ double sh[MOM_T];
    double s[MOM_NS];
#pragma unroll
    for (int k = 0; k < MOM_NS; k++) s[k] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * MOM_T + threadIdx.x; i < np; i += (int64_t)gridDim.x * MOM_T) {
        const double wi = (double)w[i];
        const double x = xh[i], y = xh[ld + i], z = xh[2 * ld + i];
        const double px = (double)u[i], py = (double)u[ld + i], pz = (double)u[2 * ld + i];
        s[0] += wi;
        s[1] += x * wi;
        s[2] += y * wi;
        s[3] += z * wi;
        s[4] += x * x * wi;
        s[5] += y * y * wi;
        s[6] += z * z * wi;
        s[7] += px * px * wi;
        s[8] += py * py * wi;
        s[9] += pz * pz * wi;
        s[10] += x * px * wi;
        s[11] += y * py * wi;
        s[12] += z * pz * wi;
    }
#pragma unroll
    for (int k = 0; k < MOM_NS; k++) {
        const double t = block_sum(s[k], sh);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * MOM_NS + k] = t;
    }
}

__global__ void __launch_bounds__(MOM_T) k_moments_final(const double* __restrict__ part, int nb,
                                                         double* __restrict__ out) {
    __shared__ double sh[MOM_T];
    for (int k = 0; k < MOM_NS; k++) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nb; b += MOM_T) v += part[(int64_t)b * MOM_NS + k];
        const double t = block_sum(v, sh);
        if (threadIdx.x == 0) out[k] = t;
    }
}

template <typename U>
static int32_t beam_moments_impl(const double* coord_half, const U* momenta, const U* wghts, int64_t ld, int32_t np,
                                 double* sums) {
    if (np < 0 || ld < np) return fail(ST_EINVAL, "beam_moments: np=%d ld=%lld", np, (long long)ld);
    if (sums == nullptr) return fail(ST_EINVAL, "beam_moments: null output");
    if (np > 0 && (!is_device_ptr(coord_half) || !is_device_ptr(momenta) || !is_device_ptr(wghts)))
        return fail(ST_EINVAL, "beam_moments: the particle arrays are device pointers");
    Stage st;
    double* d_out = st.out(sums, MOM_NS);
    CH_TRY(st.status());
    if (np == 0) {
        CH_TRY(fill_bytes(d_out, 0, MOM_NS * sizeof(double)));
        return st.finish();
    }
    const int nb = (int)std::min<int64_t>(MOM_MAXB, (np + MOM_T - 1) / MOM_T);
    double* part = static_cast<double*>(scratch(12, (size_t)nb * MOM_NS * sizeof(double)));
    if (!part) return ST_ENOMEM;
    hipLaunchKernelGGL((k_moments_part<U>), nb, MOM_T, 0, stream(), coord_half, momenta, wghts, ld, np, part);
    CH_LAUNCH_CHECK("k_moments_part");
    hipLaunchKernelGGL(k_moments_final, 1, MOM_T, 0, stream(), part, nb, d_out);
    CH_LAUNCH_CHECK("k_moments_final");
    return st.finish();
}

// ---------------------------------------------------------------- no-alloc beam moments
//
// The same two launches as beam_moments_impl, with the partial sums in the
// file-scope array instead of the library's grown scratch memory and the sums
// written straight to device memory: nothing is allocated.  The array's
// device address is looked up once per device (with that device current) and
// kept; every call on a device shares its array, so calls that overlap on one
// GPU race.  Every partial row a call reads it has written itself first, so
// nothing needs clearing.
constexpr int NOALLOC_MAX_DEVICES = 64;

static int32_t noalloc_scratch(double** out) {
    static std::mutex mu;
    static double* ptr[NOALLOC_MAX_DEVICES] = {};
    int dev = -1;
    CH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(ST_EUNSUP, "beam_moments_noalloc: device %d", dev);
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
        if (e != hipSuccess || p == nullptr)
            return fail(ST_EHIP, "beam_moments_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        ptr[dev] = static_cast<double*>(p);
    }
    *out = ptr[dev];
    return ST_OK;
}

static int32_t beam_moments_noalloc_impl(const double* coord_half, const double* momenta, const double* wghts,
                                         int64_t ld, int32_t np, double* sums) {
    if (np < 0 || ld < np) return fail(ST_EINVAL, "beam_moments_noalloc: np=%d ld=%lld", np, (long long)ld);
    if (sums == nullptr || !is_device_ptr(sums))
        return fail(ST_EINVAL, "beam_moments_noalloc: sums must be device memory (13 doubles)");
    if (np > 0 && (!is_device_ptr(coord_half) || !is_device_ptr(momenta) || !is_device_ptr(wghts)))
        return fail(ST_EINVAL, "beam_moments_noalloc: the particle arrays are device pointers");
    double* part = nullptr;
    CH_TRY(noalloc_scratch(&part));
    const int nb = (int)std::min<int64_t>(MOM_MAXB, std::max<int64_t>(1, (np + MOM_T - 1) / MOM_T));
    if ((size_t)nb * MOM_NS * sizeof(double) > sizeof(::scratch))
        return fail(ST_EINVAL, "beam_moments_noalloc: %d partial rows exceed the %d-row scratch", nb, MOM_MAXB);
    // np = 0: the grid-stride pass runs no iteration and writes zero partials
    hipLaunchKernelGGL((k_moments_part<double>), nb, MOM_T, 0, stream(), coord_half, momenta, wghts, ld, np, part);
    CH_LAUNCH_CHECK("k_moments_part");
    hipLaunchKernelGGL(k_moments_final, 1, MOM_T, 0, stream(), part, nb, sums);
    CH_LAUNCH_CHECK("k_moments_final");
    return ST_OK;
}

// ---------------------------------------------------------------- weighted mean of px
//
// PXmean of the StaticKick field update (chimera_main.py:120-121):
// sum(momenta[0]*weights) / sum(weights), both sums in FP64 by the same
// two-launch fixed-order reduction as the moments.  No particles: 0/0 = NaN,
// as the reference's NumPy quotient gives.
template <typename U>
__global__ void __launch_bounds__(MOM_T) k_pxw_part(const U* __restrict__ px, const U* __restrict__ w, int np,
                                                    double* __restrict__ part) {
    __shared__ double sh[MOM_T];
    double a = 0.0, b = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * MOM_T + threadIdx.x; i < np; i += (int64_t)gridDim.x * MOM_T) {
        const double wi = (double)w[i];
        a += (double)px[i] * wi;
        b += wi;
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) {
        part[2 * (int64_t)blockIdx.x] = a;
        part[2 * (int64_t)blockIdx.x + 1] = b;
    }
}

__global__ void __launch_bounds__(MOM_T) k_pxw_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
    __shared__ double sh[MOM_T];
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nb; k += MOM_T) {
        a += part[2 * (int64_t)k];
        b += part[2 * (int64_t)k + 1];
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) out[0] = a / b;
}

template <typename U>
static int32_t momentum_mean_impl(const U* momenta, const U* wghts, int32_t np, double* px0) {
    if (np < 0) return fail(ST_EINVAL, "momentum_mean: np=%d", np);
    if (px0 == nullptr) return fail(ST_EINVAL, "momentum_mean: null output");
    if (np > 0 && (!is_device_ptr(momenta) || !is_device_ptr(wghts)))
        return fail(ST_EINVAL, "momentum_mean: momenta and weights are device pointers");
    Stage st;
    double* d_out = st.out(px0, 1);
    CH_TRY(st.status());
    const int nb = (int)std::min<int64_t>(MOM_MAXB, std::max<int64_t>(1, (np + MOM_T - 1) / MOM_T));
    double* part = static_cast<double*>(scratch(12, (size_t)nb * 2 * sizeof(double)));
    if (!part) return ST_ENOMEM;
    hipLaunchKernelGGL((k_pxw_part<U>), nb, MOM_T, 0, stream(), momenta, wghts, np, part);
    CH_LAUNCH_CHECK("k_pxw_part");
    hipLaunchKernelGGL(k_pxw_final, 1, MOM_T, 0, stream(), part, nb, d_out);
    CH_LAUNCH_CHECK("k_pxw_final");
    return st.finish();
}

// ---------------------------------------------------------------- EM energy per kx
//
// nrg_out (diagnostics.py:109-122): out[x] = sum over k < nkr, m < nmodes of
// EnergyFact[x,k,m] * (|E_x|^2 + |E_y|^2 + |E_z|^2)[x,k,m] from EG_fb(nkx,
// nkr, nmodes, 6) (the first three components are E), summed in FP64 by one
// thread per kx in (k, m) order: the same bits on every call.
template <typename C>
__global__ void k_field_energy(const C* __restrict__ eg, const double* __restrict__ ef, int nkx, int nkr, int nmodes,
                               double* __restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nkx) return;
    const int64_t cs = (int64_t)nkx * nkr * nmodes;   // component stride
    double acc = 0.0;
    for (int m = 0; m < nmodes; m++)
        for (int k = 0; k < nkr; k++) {
            const int64_t o = x + (int64_t)nkx * (k + (int64_t)nkr * m);
            double e2 = 0.0;
            for (int c = 0; c < 3; c++) {
                const C v = eg[o + c * cs];
                e2 += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
            }
            acc += e2 * ef[o];
        }
    out[x] = acc;
}

template <typename C>
static int32_t field_energy_impl(const C* eg_fb, const double* energy_fact, int32_t nkx, int32_t nkr, int32_t nmodes,
                                 double* out) {
    if (nkx < 1 || nkr < 1 || nmodes < 1) return fail(ST_EINVAL, "field_energy: bad sizes");
    if (out == nullptr) return fail(ST_EINVAL, "field_energy: null output");
    if (!is_device_ptr(eg_fb) || !is_device_ptr(energy_fact))
        return fail(ST_EINVAL, "field_energy: EG_fb and EnergyFact are device pointers");
    Stage st;
    double* d_out = st.out(out, (size_t)nkx);
    CH_TRY(st.status());
    hipLaunchKernelGGL((k_field_energy<C>), blocks_for(nkx, 256), 256, 0, stream(), eg_fb, energy_fact, nkx, nkr,
                       nmodes, d_out);
    CH_LAUNCH_CHECK("k_field_energy");
    return st.finish();
}

// ---------------------------------------------------------------- laser power per x
//
// pwr_out (diagnostics.py:124-149) after fb_vec_out: out[x] = sum over rows
// r = 0..nr of Rgrid[r] * sum_m sum_{c<3} |vec[x,r,m,c]|^2 for vec(nkx, 0:nr,
// nmodes, >= 3); one thread per x, (r, m, c) order.
template <typename C>
__global__ void k_line_power(const C* __restrict__ vec, const double* __restrict__ rgrid, int nkx, int nr, int nmodes,
                             double* __restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nkx) return;
    const int64_t rs = nkx, ms = (int64_t)nkx * (nr + 1), cs = ms * nmodes;
    double acc = 0.0;
    for (int r = 0; r <= nr; r++) {
        double a2 = 0.0;
        for (int m = 0; m < nmodes; m++)
            for (int c = 0; c < 3; c++) {
                const C v = vec[x + r * rs + m * ms + c * cs];
                a2 += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
            }
        acc += rgrid[r] * a2;
    }
    out[x] = acc;
}

template <typename C>
static int32_t line_power_impl(const C* vec, const double* rgrid, int32_t nkx, int32_t nr, int32_t nmodes,
                               double* out) {
    if (nkx < 1 || nr < 0 || nmodes < 1) return fail(ST_EINVAL, "line_power: bad sizes");
    if (out == nullptr) return fail(ST_EINVAL, "line_power: null output");
    if (!is_device_ptr(vec) || !is_device_ptr(rgrid))
        return fail(ST_EINVAL, "line_power: vec and Rgrid are device pointers");
    Stage st;
    double* d_out = st.out(out, (size_t)nkx);
    CH_TRY(st.status());
    hipLaunchKernelGGL((k_line_power<C>), blocks_for(nkx, 256), 256, 0, stream(), vec, rgrid, nkx, nr, nmodes, d_out);
    CH_LAUNCH_CHECK("k_line_power");
    return st.finish();
}

// ---------------------------------------------------------------- phase-space dump (phs_out)
//
// The particles phs_out keeps (diagnostics.py:94-96): sum over the three
// components of (1 + u_c^2) > 20^2, in the reference's order of operations
// (each square rounded before the sum; the sum row by row).  Two launches: a
// count per chunk of SEL_CH particles, then, from the chunk offsets the host
// forms, an ordered scatter of the kept particles' half-step positions,
// momenta and weight into PartsPack(7, n) (Fortran order), so the particles
// keep their order.
constexpr int SEL_T = 256, SEL_PER = 4, SEL_CH = SEL_T * SEL_PER;

template <typename U>
__device__ __forceinline__ bool phs_keep(const U* __restrict__ u, int64_t ld, int64_t i) {
    const double a = (double)u[i], b = (double)u[ld + i], c = (double)u[2 * ld + i];
    return ((1.0 + __dmul_rn(a, a)) + (1.0 + __dmul_rn(b, b))) + (1.0 + __dmul_rn(c, c)) > 400.0;
}

// exclusive scan of one int per thread over the workgroup; total in *tot
__device__ __forceinline__ int block_excl_scan(int v, int* sh, int* tot) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 1; s < SEL_T; s <<= 1) {
        const int t = (int)threadIdx.x >= s ? sh[threadIdx.x - s] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    const int incl = sh[threadIdx.x];
    *tot = sh[SEL_T - 1];
    __syncthreads();
    return incl - v;
}

template <typename U>
__global__ void __launch_bounds__(SEL_T) k_phs_count(const U* __restrict__ u, int64_t ld, int np,
                                                     int* __restrict__ counts) {
    __shared__ int sh[SEL_T];
    const int64_t base = (int64_t)blockIdx.x * SEL_CH + (int64_t)threadIdx.x * SEL_PER;
    int c = 0;
    for (int j = 0; j < SEL_PER; j++)
        if (base + j < np && phs_keep(u, ld, base + j)) c++;
    int tot;
    block_excl_scan(c, sh, &tot);
    if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

template <typename U>
__global__ void __launch_bounds__(SEL_T) k_phs_scatter(const double* __restrict__ xh, const U* __restrict__ u,
                                                       const U* __restrict__ w, int64_t ld, int np,
                                                       const int64_t* __restrict__ offs, double* __restrict__ out) {
    __shared__ int sh[SEL_T];
    const int64_t base = (int64_t)blockIdx.x * SEL_CH + (int64_t)threadIdx.x * SEL_PER;
    bool keep[SEL_PER];
    int c = 0;
    for (int j = 0; j < SEL_PER; j++) {
        keep[j] = base + j < np && phs_keep(u, ld, base + j);
        c += keep[j];
    }
    int tot;
    int64_t o = offs[blockIdx.x] + block_excl_scan(c, sh, &tot);
    for (int j = 0; j < SEL_PER; j++) {
        if (!keep[j]) continue;
        const int64_t i = base + j;
        double* d = out + 7 * o;
        d[0] = xh[i];
        d[1] = xh[ld + i];
        d[2] = xh[2 * ld + i];
        d[3] = (double)u[i];
        d[4] = (double)u[ld + i];
        d[5] = (double)u[2 * ld + i];
        d[6] = (double)w[i];
        o++;
    }
}

template <typename U>
static int32_t phs_select_impl(const double* coord_half, const U* momenta, const U* wghts, int64_t ld, int32_t np,
                               double* out, int64_t cap, int64_t* n_sel) {
    if (np < 0 || ld < np || n_sel == nullptr) return fail(ST_EINVAL, "phs_select: np=%d ld=%lld", np, (long long)ld);
    if (np > 0 && (!is_device_ptr(coord_half) || !is_device_ptr(momenta) || !is_device_ptr(wghts)))
        return fail(ST_EINVAL, "phs_select: the particle arrays are device pointers");
    if (out != nullptr && !is_device_ptr(out)) return fail(ST_EINVAL, "phs_select: out is device memory");
    *n_sel = 0;
    if (np == 0) return ST_OK;
    const int nb = (int)((np + SEL_CH - 1) / SEL_CH);
    int* counts = static_cast<int*>(scratch(13, (size_t)nb * sizeof(int)));
    int64_t* offs = static_cast<int64_t*>(scratch(14, (size_t)nb * sizeof(int64_t)));
    if (!counts || !offs) return ST_ENOMEM;
    hipLaunchKernelGGL((k_phs_count<U>), nb, SEL_T, 0, stream(), momenta, ld, np, counts);
    CH_LAUNCH_CHECK("k_phs_count");
    std::vector<int> hc(nb);
    std::vector<int64_t> ho(nb);
    CH_HIP(hipMemcpyAsync(hc.data(), counts, (size_t)nb * sizeof(int), hipMemcpyDeviceToHost, stream()));
    CH_HIP(hipStreamSynchronize(stream()));
    int64_t tot = 0;
    for (int b = 0; b < nb; b++) {
        ho[b] = tot;
        tot += hc[b];
    }
    *n_sel = tot;
    if (out == nullptr || tot == 0) return ST_OK;
    if (tot > cap) return fail(ST_EINVAL, "phs_select: %lld particles kept, room for %lld", (long long)tot,
                               (long long)cap);
    CH_HIP(hipMemcpyAsync(offs, ho.data(), (size_t)nb * sizeof(int64_t), hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL((k_phs_scatter<U>), nb, SEL_T, 0, stream(), coord_half, momenta, wghts, ld, np, offs, out);
    CH_LAUNCH_CHECK("k_phs_scatter");
    CH_HIP(hipStreamSynchronize(stream()));     // the host offsets go out of scope
    return ST_OK;
}

// ---------------------------------------------------------------- azimuthal intensity profile (intens_profO)
//
// PWR_RO(iO, ir) = sum over l < 3 and ix of |sum_m phaseO(iO, m) Fld(ix, ir, m, l)|^2
// (utils.f90:18-57) for Fld(nkx, 0:nr, -nkO:nkO, 3) and ir = 1..nr; the phase
// table phaseO(iO, m) is formed on the host by the reference's recurrences.
// One thread per (iO, ir); the sums in the reference's (l, ix) order.
template <typename C>
__global__ void k_intens_profo(const C* __restrict__ fld, const double2* __restrict__ ph, int NO, int nkx, int nr,
                               int nm, double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NO * nr) return;
    const int iO = t % NO, ir = 1 + t / NO;
    const int64_t ms = (int64_t)nkx * (nr + 1), ls = ms * nm;
    double acc = 0.0;
    for (int l = 0; l < 3; l++)
        for (int ix = 0; ix < nkx; ix++) {
            double re = 0.0, im = 0.0;
            const C* f = fld + ix + (int64_t)nkx * ir + ls * l;
            for (int m = 0; m < nm; m++) {
                const double2 p = ph[iO + (int64_t)NO * m];
                const C v = f[ms * m];
                re += p.x * (double)v.x - p.y * (double)v.y;
                im += p.x * (double)v.y + p.y * (double)v.x;
            }
            acc += re * re + im * im;
        }
    out[iO + (int64_t)NO * (ir - 1)] = acc;
}

template <typename C>
static int32_t intens_profo_impl(double* pwr, const C* fld, int32_t NO, int32_t nkx, int32_t nr, int32_t nm) {
    if (NO < 2 || nkx < 1 || nr < 1 || nm < 1 || (nm & 1) == 0)
        return fail(ST_EINVAL, "intens_profo: NO=%d nkx=%d nr=%d and %d modes (an odd count, -nkO..nkO)", NO, nkx,
                    nr, nm);
    const int nkO = (nm - 1) / 2;
    std::vector<std::complex<double>> Os(NO), tab((size_t)NO * nm);
    const double pi = 4.0 * std::atan(1.0);
    const std::complex<double> step(std::cos(2.0 * pi / (double)(NO - 1)), std::sin(2.0 * pi / (double)(NO - 1)));
    Os[0] = 1.0;
    for (int i = 1; i < NO; i++) Os[i] = Os[i - 1] * step;
    for (int i = 0; i < NO; i++) {
        const std::complex<double> pp = Os[i], pm = 1.0 / Os[i];
        std::vector<std::complex<double>> po(nm);
        po[nkO] = 1.0;
        for (int k = 1; k <= nkO; k++) {
            po[nkO + k] = po[nkO + k - 1] * pp;
            po[nkO - k] = po[nkO - k + 1] * pm;
        }
        for (int m = 0; m < nm; m++) tab[i + (size_t)NO * m] = po[m];
    }
    Stage st;
    const C* d_fld = st.in(fld, (size_t)nkx * (nr + 1) * nm * 3);
    double* d_out = st.out(pwr, (size_t)NO * nr);
    const double2* d_ph = static_cast<const double2*>(shared_table(tab.data(), tab.size() * sizeof(double2)));
    CH_TRY(st.status());
    if (!d_ph) return ST_ENOMEM;
    hipLaunchKernelGGL((k_intens_profo<C>), blocks_for((int64_t)NO * nr, 128), 128, 0, stream(), d_fld, d_ph, NO,
                       nkx, nr, nm, d_out);
    CH_LAUNCH_CHECK("k_intens_profo");
    return st.finish();
}

}  // namespace chimera

using namespace chimera;

extern "C" {

int32_t chimera_beam_moments(const double* coord_half, const double* momenta, const double* wghts, int64_t ld,
                             int32_t np, double* sums) {
    return beam_moments_impl<double>(coord_half, momenta, wghts, ld, np, sums);
}

int32_t chimera_beam_moments_mixed(const double* coord_half, const float* momenta, const float* wghts, int64_t ld,
                                   int32_t np, double* sums) {
    return beam_moments_impl<float>(coord_half, momenta, wghts, ld, np, sums);
}

int32_t chimera_beam_moments_noalloc(const double* coord_half, const double* momenta, const double* wghts,
                                     int64_t ld, int32_t np, double* sums) {
    return beam_moments_noalloc_impl(coord_half, momenta, wghts, ld, np, sums);
}

int32_t chimera_momentum_mean(const double* momenta, const double* wghts, int32_t np, double* px0) {
    return momentum_mean_impl<double>(momenta, wghts, np, px0);
}

int32_t chimera_momentum_mean_mixed(const float* momenta, const float* wghts, int32_t np, double* px0) {
    return momentum_mean_impl<float>(momenta, wghts, np, px0);
}

int32_t chimera_field_energy(const double* EG_fb, const double* energy_fact, int32_t nkx, int32_t nkr, int32_t nmodes,
                             double* out) {
    return field_energy_impl<double2>(reinterpret_cast<const double2*>(EG_fb), energy_fact, nkx, nkr, nmodes, out);
}

int32_t chimera_field_energy_c8(const float* EG_fb, const double* energy_fact, int32_t nkx, int32_t nkr,
                                int32_t nmodes, double* out) {
    return field_energy_impl<float2>(reinterpret_cast<const float2*>(EG_fb), energy_fact, nkx, nkr, nmodes, out);
}

int32_t chimera_line_power(const double* vec, const double* Rgrid, int32_t nkx, int32_t nr, int32_t nmodes,
                           double* out) {
    return line_power_impl<double2>(reinterpret_cast<const double2*>(vec), Rgrid, nkx, nr, nmodes, out);
}

int32_t chimera_line_power_c8(const float* vec, const double* Rgrid, int32_t nkx, int32_t nr, int32_t nmodes,
                              double* out) {
    return line_power_impl<float2>(reinterpret_cast<const float2*>(vec), Rgrid, nkx, nr, nmodes, out);
}

int32_t chimera_phs_select(const double* coord_half, const double* momenta, const double* wghts, int64_t ld,
                           int32_t np, double* out, int64_t cap, int64_t* n_sel) {
    return phs_select_impl<double>(coord_half, momenta, wghts, ld, np, out, cap, n_sel);
}

int32_t chimera_phs_select_mixed(const double* coord_half, const float* momenta, const float* wghts, int64_t ld,
                                 int32_t np, double* out, int64_t cap, int64_t* n_sel) {
    return phs_select_impl<float>(coord_half, momenta, wghts, ld, np, out, cap, n_sel);
}

int32_t chimera_intens_profo(double* PWR_RO, const double* Fld, int32_t NO, int32_t nkx, int32_t nr, int32_t nmodes) {
    return intens_profo_impl<double2>(PWR_RO, reinterpret_cast<const double2*>(Fld), NO, nkx, nr, nmodes);
}

int32_t chimera_intens_profo_c8(double* PWR_RO, const float* Fld, int32_t NO, int32_t nkx, int32_t nr,
                                int32_t nmodes) {
    return intens_profo_impl<float2>(PWR_RO, reinterpret_cast<const float2*>(Fld), NO, nkx, nr, nmodes);
}

}  // extern "C"
