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

#include "common.hpp"
#include "../../include/chimera_hip.h"

namespace chimera {

constexpr int MOM_T = 256, MOM_NS = 13, MOM_MAXB = 512;

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

}  // extern "C"
