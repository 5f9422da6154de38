// fb.hip -- the Fourier-Bessel side of the step: the forward and inverse FB
// transforms of fb_io.f90 (discrete Hankel transform along r + FFT along x),
// the absorbing-layer filter, and the FB-space curl / grad / div / grad-div of
// fb_math.f90.
//
// Every dense contraction (Hankel matrices In/Out, the mode-coupling matrices
// DpS2S/DmS2S) goes through ONE job-driven kernel, k_fb_gemm: a job is an
// output mode of one operator, made of up to two K-segments
//   P_s(x,k) = sum_k' A_s(x,k') D_s(k',k),   A_s = c0*T(a0) + c1*T(a1)
// (T = identity, or the reflected conjugate -conj(v(-kx)) that stands for mode
// -1 at m = 0) accumulated into one or two accumulators, and up to three
// outputs, each a complex combination of the accumulators plus the diagonal
// i*kx term of the operator.  All jobs of one operator run in one launch
// (grid.z = job), and no two jobs write the same output.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <string>
#include <complex>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "fb_kernels.hpp"
#include "fft.hpp"
#include "../../include/chimera_hip.h"

namespace chimera {

// ================================================================== FFT

const double2* fft_twiddles(int N) {
    static std::mutex mu;
    static std::map<int, const double2*> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(N);
    if (it != cache.end()) return it->second;
    std::vector<double2> h(N > 0 ? N : 1);
    for (int k = 0; k < N; k++) {
        // exact argument reduction: angle = -2 pi k/N, symmetric in k and N-k
        const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)N;
        h[k] = make_double2((double)cosl(a), (double)sinl(a));
    }
    const double2* d = static_cast<const double2*>(shared_table(h.data(), h.size() * sizeof(double2)));
    if (d) cache[N] = d;
    return d;
}

const float2* fft_twiddles_f(int N) {
    static std::mutex mu;
    static std::map<int, const float2*> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(N);
    if (it != cache.end()) return it->second;
    std::vector<float2> h(N > 0 ? N : 1);
    for (int k = 0; k < N; k++) {
        const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)N;
        h[k] = make_float2((float)cosl(a), (float)sinl(a));
    }
    const float2* d = static_cast<const float2*>(shared_table(h.data(), h.size() * sizeof(float2)));
    if (d) cache[N] = d;
    return d;
}

// Store of one transformed point: post-multiplier, real scale table, mode
// normalisation and the ghost row (shared by both line kernels).
template <typename R>
__device__ __forceinline__ void fft_store(const FftArgsT<R>& a, int64_t line, int x, cplx_t<R> v) {
    typedef cplx_t<R> C;
    const int64_t r = line % a.n_inner, q = line / a.n_inner;
    if (a.post) v = cmul(v, a.post[x]);
    if (a.scale) v = crs(a.scale[x + (int64_t)a.N * r], v);
    int m = 0;
    if (a.ghost) {
        m = (int)(q % a.nmodes);
        if (a.ghost == 2) v = crs((R)(m == 0 ? 0.5 / M_PI : 1.0 / M_PI), v);
        else if (a.ghost >= 3) v = crs((R)(1.0 / M_PI), v);
    }
    C* o = a.out + r * a.out_r + q * a.out_q + x;
    *o = v;
    if (a.ghost && r == 0) {
        // ghost radial row below row 1: zero (fb_io.f90:199), +-row 1 by mode
        // (grid_deps.f90:248-265), or +row 1 (3) / -row 1 (4) for every mode of
        // the envelope grid (grid_deps_env.f90:240-283: + when nkO = 0)
        o[-a.out_r] = a.ghost == 1 ? cmake<R>(0, 0)
                    : a.ghost == 2 ? (m == 0 ? v : cneg(v))
                    : a.ghost == 3 ? v : cneg(v);
    }
}

template <int EPT, typename R, int TPB = FFT_THREADS>
__global__ void __launch_bounds__(TPB) k_fft_lines(FftArgsT<R> a) {
    typedef cplx_t<R> C;
    constexpr int TOT = EPT * TPB;
    extern __shared__ __align__(16) unsigned char fft_smem[];
    C* const sbuf = reinterpret_cast<C*>(fft_smem);
    const int N = a.N;
    const int lpb = TOT / N;
    const int64_t line0 = (int64_t)blockIdx.x * lpb;
    const C* tw = a.tw;
    const C* sph = nullptr;
    if constexpr (EPT == 8) {
        // lines of up to 2048 points: the loads go to registers first; where
        // the LDS has room the twiddle table is staged there (the passes read
        // it there instead of waiting on memory once per pass) and the phase
        // table formed there while the loads are in flight
        C vin[EPT];
#pragma unroll
        for (int j = 0; j < EPT; j++) {
            const int i = threadIdx.x + TPB * j;
            const int64_t line = line0 + i / N;
            vin[j] = cmake<R>(0, 0);
            if (line < a.nlines) {
                const int64_t r = line % a.n_inner, q = line / a.n_inner;
                vin[j] = a.in[r * a.in_r + q * a.in_q + i % N];
            }
        }
        C* const stw = sbuf + TOT;
        if (a.tw_lds) {
            for (int i = threadIdx.x; i < N; i += TPB) stw[i] = a.tw[i];
            tw = stw;
        }
        if (a.ph_at) {
            // the table of k_phase, formed here: one launch fewer per transform
            C* const p = stw + N;
            const double lx = a.ph_leftX_dev ? *a.ph_leftX_dev : a.ph_leftX;
            for (int i = threadIdx.x; i < N; i += TPB) {
                double sn, cs;
                sincos(lx * a.ph_kx[i], &sn, &cs);
                p[i] = cmake<R>((R)cs, (R)(a.ph_s * sn));
            }
            sph = p;
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < EPT; j++) {
            const int i = threadIdx.x + TPB * j;
            const int x = i % N;
            C v = vin[j];
            if (line0 + i / N < a.nlines) {
                if (a.pre) v = cmul(v, a.pre[x]);
                else if (a.ph_at == 1) v = cmul(v, sph[x]);
            }
            sbuf[i] = v;
        }
    } else {
        // ---- load (with the per-x pre-multiplier)
        for (int i = threadIdx.x; i < TOT; i += TPB) {
            const int64_t line = line0 + i / N;
            const int x = i % N;
            C v = cmake<R>(0, 0);
            if (line < a.nlines) {
                const int64_t r = line % a.n_inner, q = line / a.n_inner;
                v = a.in[r * a.in_r + q * a.in_q + x];
                if (a.pre) v = cmul(v, a.pre[x]);
            }
            sbuf[i] = v;
        }
    }
    __syncthreads();
    fft_lds<EPT, R, TPB>(sbuf, N, a.log2N, a.sign, tw);
    if (a.filtr) {
        // fb_io.f90:263: multiply the first nxfilt points, transform back
        for (int i = threadIdx.x; i < TOT; i += TPB) {
            const int x = i % N;
            if (x < a.nxfilt) sbuf[i] = crs(a.filtr[x], sbuf[i]);
        }
        __syncthreads();
        fft_lds<EPT, R, TPB>(sbuf, N, a.log2N, -a.sign, tw);
    }
    // ---- store (with post-multiplier, real scale table, mode normalisation)
    for (int i = threadIdx.x; i < TOT; i += TPB) {
        const int64_t line = line0 + i / N;
        if (line >= a.nlines) continue;
        const int x = i % N;
        C v = sbuf[i];
        if (a.ph_at == 2) v = cmul(v, sph[x]);
        fft_store<R>(a, line, x, v);
    }
}

// ---- lengths that are not powers of two (Bluestein)
// A length-N DFT of sign s is X_k = c_k * sum_n (x_n c_n) conj(c_(k-n)) with the
// chirp c_n = exp(s i pi n^2 / N) (from nk = (n^2 + k^2 - (k-n)^2) / 2): a
// circular convolution, of power-of-two length L >= 2N - 1, of the chirped,
// zero-padded line with g, g[m] = conj(c_|m|) for |m| < N.  ghat = FFT_L(g) / L,
// so that IFFT_L(FFT_L(a) * ghat) (both unnormalised) is the convolution.  The
// tables are built in long double with the angle pi (n^2 mod 2N) / N reduced
// exactly, then rounded to R.
static void host_fft_ld(std::vector<std::complex<long double>>& v) {
    // iterative radix-2, forward (exp(-2 pi i jk / L)), in place
    const size_t L = v.size();
    for (size_t i = 1, j = 0; i < L; i++) {
        size_t bit = L >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(v[i], v[j]);
    }
    const long double pi = 3.141592653589793238462643383279502884L;
    for (size_t len = 2; len <= L; len <<= 1) {
        for (size_t i = 0; i < L; i += len) {
            for (size_t k = 0; k < len / 2; k++) {
                const long double ang = -2.0L * pi * (long double)k / (long double)len;
                const std::complex<long double> w(cosl(ang), sinl(ang));
                const std::complex<long double> u = v[i + k], t = w * v[i + k + len / 2];
                v[i + k] = u + t;
                v[i + k + len / 2] = u - t;
            }
        }
    }
}

int bluestein_length(int N) {
    int L = 1;
    while (L < 2 * N - 1) L <<= 1;
    return L;
}

template <typename R>
static int32_t bluestein_tables(int N, int sign, const cplx_t<R>** chirp, const cplx_t<R>** ghat) {
    typedef cplx_t<R> C;
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::pair<const C*, const C*>> cache;
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_pair(N, sign);
    auto it = cache.find(key);
    if (it == cache.end()) {
        const int L = bluestein_length(N);
        const long double pi = 3.141592653589793238462643383279502884L;
        std::vector<std::complex<long double>> c(N), g(L, std::complex<long double>(0, 0));
        for (int n = 0; n < N; n++) {
            const long long n2 = ((long long)n * n) % (2LL * N);
            const long double ang = (long double)sign * pi * (long double)n2 / (long double)N;
            c[n] = std::complex<long double>(cosl(ang), sinl(ang));
        }
        g[0] = std::conj(c[0]);
        for (int m = 1; m < N; m++) g[m] = g[L - m] = std::conj(c[m]);
        host_fft_ld(g);
        std::vector<C> hc(N), hg(L);
        for (int n = 0; n < N; n++) hc[n] = cmake<R>((R)c[n].real(), (R)c[n].imag());
        for (int k = 0; k < L; k++) hg[k] = cmake<R>((R)(g[k].real() / L), (R)(g[k].imag() / L));
        const C* dc = static_cast<const C*>(shared_table(hc.data(), hc.size() * sizeof(C)));
        const C* dg = static_cast<const C*>(shared_table(hg.data(), hg.size() * sizeof(C)));
        if (!dc || !dg) return ST_ENOMEM;
        it = cache.emplace(key, std::make_pair(dc, dg)).first;
    }
    *chirp = it->second.first;
    *ghat = it->second.second;
    return ST_OK;
}

// The convolution step on the lines of buf (stride L): forward FFT, times ghat,
// unnormalised backward FFT.
template <int EPT, typename R>
__device__ __forceinline__ void bluestein_conv(cplx_t<R>* buf, int L, int log2L, const cplx_t<R>* __restrict__ ghat,
                                               const cplx_t<R>* __restrict__ tw) {
    constexpr int TOT = EPT * FFT_THREADS;
    fft_lds<EPT, R>(buf, L, log2L, -1, tw);
    for (int i = threadIdx.x; i < TOT; i += FFT_THREADS) buf[i] = cmul(buf[i], ghat[i & (L - 1)]);
    __syncthreads();
    fft_lds<EPT, R>(buf, L, log2L, +1, tw);
}

// Lines of length N (not a power of two, 2N - 1 <= 8192) through Bluestein:
// TOT / L lines per block, each in L points of LDS.  Same loads, pre/post
// multipliers, absorbing-layer pass and stores as k_fft_lines.
template <int EPT, typename R>
__global__ void __launch_bounds__(FFT_THREADS) k_fft_bluestein(FftArgsT<R> a) {
    typedef cplx_t<R> C;
    constexpr int TOT = EPT * FFT_THREADS;
    extern __shared__ __align__(16) unsigned char fft_smem[];
    C* const sbuf = reinterpret_cast<C*>(fft_smem);
    const int N = a.N, lm = a.L - 1;
    const int64_t line0 = (int64_t)blockIdx.x * (TOT >> a.log2L);
    for (int i = threadIdx.x; i < TOT; i += FFT_THREADS) {
        const int64_t line = line0 + (i >> a.log2L);
        const int x = i & lm;
        C v = cmake<R>(0, 0);
        if (x < N && line < a.nlines) {
            const int64_t r = line % a.n_inner, q = line / a.n_inner;
            v = a.in[r * a.in_r + q * a.in_q + x];
            if (a.pre) v = cmul(v, a.pre[x]);
            v = cmul(v, a.bs_chirp[x]);
        }
        sbuf[i] = v;
    }
    __syncthreads();
    bluestein_conv<EPT, R>(sbuf, a.L, a.log2L, a.bs_ghat, a.tw);
    const C* chirp = a.bs_chirp;
    if (a.filtr) {
        // fb_io.f90:263: multiply the first nxfilt points, transform back
        for (int i = threadIdx.x; i < TOT; i += FFT_THREADS) {
            const int x = i & lm;
            C v = cmake<R>(0, 0);
            if (x < N) {
                v = cmul(sbuf[i], a.bs_chirp[x]);
                if (x < a.nxfilt) v = crs(a.filtr[x], v);
                v = cmul(v, a.bs_chirp2[x]);
            }
            sbuf[i] = v;
        }
        __syncthreads();
        bluestein_conv<EPT, R>(sbuf, a.L, a.log2L, a.bs_ghat2, a.tw);
        chirp = a.bs_chirp2;
    }
    for (int i = threadIdx.x; i < TOT; i += FFT_THREADS) {
        const int64_t line = line0 + (i >> a.log2L);
        const int x = i & lm;
        if (x >= N || line >= a.nlines) continue;
        fft_store<R>(a, line, x, cmul(sbuf[i], chirp[x]));
    }
}

int32_t fft_length_check(int N) {
    if (N < 1 || N > 8192 || ((N & (N - 1)) != 0 && bluestein_length(N) > 8192))
        return fail(ST_EUNSUP,
                    "FFT length %d: this build transforms powers of two up to 8192 and other lengths up to 4096 "
                    "along x", N);
    return ST_OK;
}

template <typename R>
static int32_t launch_fft_bluestein(FftArgsT<R> a) {
    typedef cplx_t<R> C;
    const int N = a.N, L = bluestein_length(N);
    a.L = L;
    a.log2L = 0;
    while ((1 << a.log2L) < L) a.log2L++;
    if constexpr (sizeof(R) == 8) a.tw = fft_twiddles(L);
    else a.tw = fft_twiddles_f(L);
    if (!a.tw) return ST_ENOMEM;
    CH_TRY(bluestein_tables<R>(N, a.sign, &a.bs_chirp, &a.bs_ghat));
    if (a.filtr) CH_TRY(bluestein_tables<R>(N, -a.sign, &a.bs_chirp2, &a.bs_ghat2));
    a.tw_lds = 0;
    if (a.ph_at) {
        CH_TRY(make_phase<R>(a.ph_table, a.ph_kx, a.ph_leftX, a.ph_s, N, a.ph_leftX_dev));
        if (a.ph_at == 1) a.pre = a.ph_table;
        else a.post = a.ph_table;
        a.ph_at = 0;
    }
    const int ept = L <= 2048 ? 8 : L / FFT_THREADS;
    const int tot = ept * FFT_THREADS;
    const unsigned grid = (unsigned)((a.nlines + tot / L - 1) / (tot / L));
    const size_t shmem = (size_t)tot * sizeof(C);
    static bool attr_done = false;
    if (!attr_done) {
        CH_HIP(hipFuncSetAttribute((const void*)k_fft_bluestein<8, R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(8 * FFT_THREADS * sizeof(C))));
        CH_HIP(hipFuncSetAttribute((const void*)k_fft_bluestein<16, R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(16 * FFT_THREADS * sizeof(C))));
        CH_HIP(hipFuncSetAttribute((const void*)k_fft_bluestein<32, R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(32 * FFT_THREADS * sizeof(C))));
        attr_done = true;
    }
    KTimer kt(KF_FFT);
    switch (ept) {
        case 8:
            hipLaunchKernelGGL((k_fft_bluestein<8, R>), grid, FFT_THREADS, shmem, stream(), a);
            break;
        case 16:
            hipLaunchKernelGGL((k_fft_bluestein<16, R>), grid, FFT_THREADS, shmem, stream(), a);
            break;
        case 32:
            hipLaunchKernelGGL((k_fft_bluestein<32, R>), grid, FFT_THREADS, shmem, stream(), a);
            break;
        default:
            return fail(ST_EUNSUP, "FFT length %d", N);
    }
    CH_LAUNCH_CHECK("k_fft_bluestein");
    return ST_OK;
}

// Threads per workgroup for lines of 4096 and 8192 points, 8 points each.
constexpr int FFT_4K_TPB = 4096 / 8, FFT_8K_TPB = 8192 / 8;

template <typename R>
int32_t launch_fft(FftArgsT<R> a) {
    typedef cplx_t<R> C;
    if (a.nlines <= 0) return ST_OK;
    const int N = a.N;
    CH_TRY(fft_length_check(N));
    if ((N & (N - 1)) != 0) return launch_fft_bluestein<R>(a);
    int l2 = 0;
    while ((1 << l2) < N) l2++;
    a.log2N = l2;
    if constexpr (sizeof(R) == 8) a.tw = fft_twiddles(N);
    else a.tw = fft_twiddles_f(N);
    if (!a.tw) return ST_ENOMEM;
    // every line length runs 8 points a thread (radix-8 passes): lines of up
    // to 2048 points with 256 threads, several lines per workgroup; longer
    // lines fill a workgroup's LDS alone (64 and 128 KiB in complex128) with
    // N / 8 threads, 8 and 16 waves, to hide the passes' LDS latency
    const int tpb = N <= 2048 ? FFT_THREADS : N / 8;
    const int ept = 8;
    const int tot = ept * tpb;
    const int lpb = tot / N;
    const unsigned grid = (unsigned)((a.nlines + lpb - 1) / lpb);
    // LDS of an EPT = 8 workgroup kept within 48 KiB (three per CU): the
    // twiddles and the phase table are staged there only when they fit
    a.tw_lds = ept == 8 && (size_t)(tot + N) * sizeof(C) <= 48 * 1024;
    if (a.ph_at && !(a.tw_lds && (size_t)(tot + 2 * N) * sizeof(C) <= 48 * 1024)) {
        // no LDS room for the phase table in the transform: k_phase forms it
        CH_TRY(make_phase<R>(a.ph_table, a.ph_kx, a.ph_leftX, a.ph_s, N, a.ph_leftX_dev));
        if (a.ph_at == 1) a.pre = a.ph_table;
        else a.post = a.ph_table;
        a.ph_at = 0;
    }
    const size_t shmem = (size_t)(tot + (a.tw_lds ? N : 0) + (a.ph_at ? N : 0)) * sizeof(C);
    static bool attr_done = false;
    if (!attr_done) {
        // complex128 lines of 4096 and 8192 points use 64 and 128 KiB of LDS per
        // workgroup; EPT = 8 at most 64 KiB (2048 complex128 points and their twiddles)
        CH_HIP(hipFuncSetAttribute((const void*)k_fft_lines<8, R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(64 * 1024)));
        CH_HIP(hipFuncSetAttribute((const void*)k_fft_lines<8, R, FFT_4K_TPB>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)(4096 * sizeof(C))));
        CH_HIP(hipFuncSetAttribute((const void*)k_fft_lines<8, R, FFT_8K_TPB>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)(8192 * sizeof(C))));
        attr_done = true;
    }
    KTimer kt(KF_FFT);
    if (tpb == FFT_THREADS) {
        hipLaunchKernelGGL((k_fft_lines<8, R>), grid, FFT_THREADS, shmem, stream(), a);
    } else if (tpb == FFT_4K_TPB) {
        hipLaunchKernelGGL((k_fft_lines<8, R, FFT_4K_TPB>), grid, FFT_4K_TPB, shmem, stream(), a);
    } else if (tpb == FFT_8K_TPB) {
        hipLaunchKernelGGL((k_fft_lines<8, R, FFT_8K_TPB>), grid, FFT_8K_TPB, shmem, stream(), a);
    } else {
        return fail(ST_EUNSUP, "FFT length %d", N);
    }
    CH_LAUNCH_CHECK("k_fft_lines");
    return ST_OK;
}
template int32_t launch_fft<double>(FftArgsT<double>);
template int32_t launch_fft<float>(FftArgsT<float>);

// Per-x phase exp(s*i*kx*leftX) (fb_io.f90:40, :120, :202), on the device.
template <typename R>
__global__ void k_phase(cplx_t<R>* __restrict__ ph, const double* __restrict__ kx, double leftX,
                        const double* __restrict__ leftX_dev, double s, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double lx = leftX_dev ? *leftX_dev : leftX;
    double sn, cs;
    sincos(lx * kx[i], &sn, &cs);
    ph[i] = cmake<R>((R)cs, (R)(s * sn));
}

// 1/(N*exp(i*kx*leftX)) as the reference forms it (fb_io.f90:247).
template <typename R>
__global__ void k_phase_inv(cplx_t<R>* __restrict__ ph, const double* __restrict__ kx, double leftX, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sn, cs;
    sincos(leftX * kx[i], &sn, &cs);
    const double a = (double)n * cs, b = (double)n * sn;
    const double d = a * a + b * b;
    ph[i] = cmake<R>((R)(a / d), (R)(-b / d));
}

template <typename R>
int32_t make_phase(cplx_t<R>* ph, const double* kx, double leftX, double s, int n, const double* leftX_dev) {
    hipLaunchKernelGGL((k_phase<R>), blocks_for(n, 256), 256, 0, stream(), ph, kx, leftX, leftX_dev, s, n);
    CH_LAUNCH_CHECK("k_phase");
    return ST_OK;
}
template int32_t make_phase<double>(double2*, const double*, double, double, int, const double*);
template int32_t make_phase<float>(float2*, const double*, double, double, int, const double*);

// ================================================================== FB GEMM

// Block tile: GX complex x (2*GX real rows; GX = GB_X, or GB_X/2 or GB_X/4
// for the smallest launches) by 32*WN output columns; K staged GB_K at a time through
// LDS.  Four waves in a 2 x 2 arrangement, each owning GX real rows x 16*WN
// columns as (GX/16) x WN tiles of v_mfma_f64_16x16x4f64.
// Lane layout of that instruction on gfx950 (checked on the device): lane l
// supplies A[l%16][l/16] and B[l/16][l%16] and holds D[l/16 + 4i][l%16],
// i = 0..3.  Each 16-row MFMA tile covers 8 complex x; complex c (0..7) puts
// its real part in row (c%4) + 8*(c/4) and its imaginary part 4 rows below, so
// that a lane's items (0,1) and (2,3) are the (re, im) of x = c and x = c + 4
// with c = l/16.
constexpr int GB_X = 64, GB_K = 16, GB_THREADS = 256;
// LDS row padding in doubles: rows 16 doubles (128 B) apart modulo 256 B put the
// two rows that a half-wave reads in one ds_read_b64 on disjoint bank halves.
constexpr int GB_PAD = 16;
typedef double dbl4 __attribute__((ext_vector_type(4)));

// Element (x, k') of segment s reads a0/a1 at seg_off; seg_finish turns the raw
// loads into c0*T(a0) + c1*T(a1).  Loads and arithmetic are split so that the
// loads of the next K-stage stay in flight while the MFMAs run.
template <typename R>
__device__ __forceinline__ int64_t seg_off(const FbSegT<R>& s, int64_t x, int64_t kk, int nx, int64_t lda) {
    const int64_t xr = s.refl ? (x == 0 ? 0 : nx - x) : x;
    return xr + lda * kk;
}

__device__ __forceinline__ double2 seg_finish(const FbSeg& s, double2 v0, double2 v1) {
    if (s.refl) {
        v0 = cneg(cconj(v0));
        v1 = cneg(cconj(v1));
    }
    double2 v = cmul(s.c0, v0);
    if (s.a1) v = cadd(v, cmul(s.c1, v1));
    return v;
}

__device__ __forceinline__ float2 seg_finish(const FbSegT<float>& s, float2 v0, float2 v1) {
    if (s.refl) {
        v0 = cneg(cconj(v0));
        v1 = cneg(cconj(v1));
    }
    float2 v = cmul(to_f(s.c0), v0);
    if (s.a1) v = cadd(v, cmul(to_f(s.c1), v1));
    return v;
}

// D stage in LDS: column-major, column n at n*GB_LDK, GB_LDK = GB_K + 1 doubles
// (odd), so that the 16 columns a quarter-wave reads for one k, and the two k
// of a half-wave, fall on distinct banks, and the 16 k of one column that a
// quarter-wave stores are contiguous.  (Row-major rows padded to 144 doubles
// put the stores of 16 consecutive k on two bank pairs: an 8-way conflict on
// every stage.)
constexpr int GB_LDK = GB_K + 1;

// One K-stage (GB_K = 16) of a wave's (16*NI) x (16*WN) tile: NI x WN MFMA
// tiles (NI = 4 for 64-complex x tiles, 2 for 32-complex ones).
template <int WN, int NI>
__device__ __forceinline__ void mfma_stage(dbl4 (&acc)[NI][WN], const double* __restrict__ As,
                                           const double* __restrict__ Ds, int lane, int wr, int wc) {
    constexpr int LDA = 32 * NI + GB_PAD;
#pragma unroll
    for (int kq = 0; kq < GB_K / 4; kq++) {
        const int kk = kq * 4 + (lane >> 4);
        double a[NI], b[WN];
#pragma unroll
        for (int i = 0; i < NI; i++) a[i] = As[kk * LDA + wr * 16 * NI + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < WN; j++) b[j] = Ds[(wc * 16 * WN + j * 16 + (lane & 15)) * GB_LDK + kk];
#pragma unroll
        for (int i = 0; i < NI; i++)
#pragma unroll
            for (int j = 0; j < WN; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
}

// LDS bytes of one k_fb_gemm<NACC, WN, GX> block: A and D stages, double-buffered.
constexpr size_t gb_lds_bytes(int WN, int GX) {
    return 2 * ((size_t)GB_K * (2 * GX + GB_PAD) + (size_t)32 * WN * GB_LDK) * sizeof(double);
}

// Fast-path loads of the K-stages.  With the thread layout of the fetch
// (element e = tid + GB_THREADS*i), a thread's A elements of a stage are
// (x0 + tid%GX, k0 + tid/GX + (GB_THREADS/GX)*i) and its D elements
// (k0 + tid%KD, n0 + tid/KD + (GB_THREADS/KD)*i): a per-lane byte offset that
// is the same for every stage plus a uniform one.  They are issued as buffer
// loads (segment base in a buffer resource, lane offset in a VGPR, uniform
// offset in an SGPR) instead of a 64-bit address computation per element.
// Used when the whole stage is in range (full x tile, full column tile, no K
// remainder) and the host has checked that every offset of the launch fits 31
// bits (FbGemmArgs::buf_ok); the other stages take the guarded form.
typedef float ch_vf2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t ch_rsrc(const void* p) {
    // raw buffer (stride 0), the whole 32-bit offset range, gfx9 word 3
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)0xFFFFFFFFu, 0x00020000);
}
// (the loaded vector is reinterpreted whole: taking its dwords one by one into
// floats let the compiler narrow the 64-bit load to its first dword)
__device__ __forceinline__ double2 ldb_c(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, double2) {
    const ch_vd2 v = __builtin_bit_cast(ch_vd2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 ldb_c(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float2) {
    const ch_vf2 v = __builtin_bit_cast(ch_vf2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ double ldb_r(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ float ldb_r(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
// Lane byte offset of a thread's A elements in segment S (x reflected for refl).
template <typename R>
__device__ __forceinline__ uint32_t seg_lane_a(const FbSegT<R>& S, int xg, int nx, int64_t lda, int kq0) {
    const int xc = xg < nx ? xg : 0;
    const int xr = S.refl ? (xc == 0 ? 0 : nx - xc) : xc;
    return (uint32_t)((xr + lda * kq0) * (int64_t)sizeof(cplx_t<R>));
}
// c0*T(a0) + c1*T(a1) is a0 itself (bit for bit, for finite values)
template <typename R>
__device__ __forceinline__ bool seg_plain(const FbSegT<R>& S) {
    return !S.refl && !S.a1 && S.c0.x == 1.0 && S.c0.y == 0.0;
}

// Output elements per batch of the epilogue, by element type and accumulator
// registers (a batch holds up to seven loaded values and an offset per element).
constexpr int gemm_eb(int rbytes, int nacc, int wn) {
    return rbytes == 4 ? (nacc * wn <= 2 ? 8 : nacc * wn == 4 ? 4 : 2) : (nacc * wn <= 2 ? 4 : wn == 4 ? 1 : 2);
}

// Epilogue of both GEMM kernels.  For row tile i, column tile j and h = 0, 1 a
// lane holds the complex output at column n0 + wc*16*WN + j*16 + lane%16 and
// x = x0 + wr*32 + i*8 + (FP64: lane/16 + 4h; FP32: 2*(lane/16) + h).  The
// output descriptor is copied to registers (read through the job pointer after
// a store, it would be read again, since the outputs could alias the job
// table), and the elements go in batches of EB whose loads are all issued
// before the batch's stores: one memory round trip per batch instead of
// several per element.  Every element is read and written by its own lane
// only, so the results are those of one element at a time.
template <typename R, int NACC, int WN, int NI, typename ACC>
__device__ __forceinline__ void gemm_epilogue_serial(const FbJobT<R>& J, const FbGemmArgs& g, const ACC (&acc0)[NI][WN],
                                                     const ACC (&acc1)[NI][NACC == 2 ? WN : 1], int Ncol, int n0,
                                                     int x0, int lane, int wr, int wc);

template <typename R, int NACC, int WN, int NI, typename ACC>
__device__ __forceinline__ void gemm_epilogue(const FbJobT<R>& J, const FbGemmArgs& g, const ACC (&acc0)[NI][WN],
                                              const ACC (&acc1)[NI][NACC == 2 ? WN : 1], int Ncol, int n0, int x0,
                                              int lane, int wr, int wc) {
    if constexpr (WN == 4) {
        // the 128-column tiles of large grids have no registers to spare for a
        // batch; their launches have enough blocks to hide the latency
        gemm_epilogue_serial<R, NACC, WN, NI>(J, g, acc0, acc1, Ncol, n0, x0, lane, wr, wc);
        return;
    }
    typedef cplx_t<R> C;
    constexpr int NE = NI * WN * 2, EB0 = gemm_eb((int)sizeof(R), NACC, WN), EB = EB0 < NE ? EB0 : NE;
    static_assert(NE % EB == 0, "epilogue batch");
    const int nout = J.nout;
    const bool beta = J.beta != 0;
    for (int o = 0; o < nout; o++) {
        const FbOutT<R> O = J.out[o];
        const C al0 = cconv<R>(O.alpha0), al1 = cconv<R>(O.alpha1), gam = cconv<R>(O.gamma);
        const R dti = (R)g.dt_inv;
        const bool rd_p = O.poiss || beta;
#pragma unroll
        for (int b = 0; b < NE; b += EB) {
            // out-of-range elements load element 0 (a valid address) and store nothing,
            // so that the loads need no divergent branches
            int64_t off[EB];
            int xs[EB];
            bool ok[EB];
            C v[EB], dg[EB], pv[EB], n1[EB], n0v[EB];
            R sc[EB], w2[EB];
            double kxv[EB];
#pragma unroll
            for (int e = 0; e < EB; e++) {
                const int q = b + e, i = q / (2 * WN), j = (q / 2) % WN, h = q % 2;
                const int64_t ng = n0 + wc * 16 * WN + j * 16 + (lane & 15);
                const int xg = x0 + wr * 8 * NI + i * 8 + (sizeof(R) == 8 ? (lane >> 4) + 4 * h : 2 * (lane >> 4) + h);
                ok[e] = ng < Ncol && xg < g.nx;
                off[e] = ok[e] ? xg + g.ldo * ng : 0;
                xs[e] = ok[e] ? xg : 0;
                v[e] = cmul(al0, cmake<R>(acc0[i][j][2 * h], acc0[i][j][2 * h + 1]));
                if constexpr (NACC == 2) v[e] = cadd(v[e], cmul(al1, cmake<R>(acc1[i][j][2 * h], acc1[i][j][2 * h + 1])));
            }
            if (O.diag) {
#pragma unroll
                for (int e = 0; e < EB; e++) {
                    dg[e] = gld2(O.diag, off[e]);
                    kxv[e] = g.kx[xs[e]];
                }
            }
            if (O.scale) {
#pragma unroll
                for (int e = 0; e < EB; e++) sc[e] = gld1(O.scale, off[e]);
            }
            if (O.poiss) {
#pragma unroll
                for (int e = 0; e < EB; e++) {
                    n1[e] = gld2(O.poiss_np1, off[e]);
                    n0v[e] = gld2(O.poiss_n, off[e]);
                    w2[e] = gld1(O.poiss_w2, off[e]);
                }
            }
            if (rd_p) {
#pragma unroll
                for (int e = 0; e < EB; e++) pv[e] = gld2(O.p, off[e]);
            }
#pragma unroll
            for (int e = 0; e < EB; e++) {
                C w = v[e];
                if (O.diag) w = cadd(w, crs((R)kxv[e], cmul(gam, dg[e])));
                if (O.scale) w = crs(sc[e], w);
                if (O.poiss) {
                    // maxwell_solvers.f90:153-158: J += (v + (g_np1 - g_n)*dt_inv)*w2_inv
                    w = cadd(pv[e], crs(w2[e], cadd(w, crs(dti, csub(n1[e], n0v[e])))));
                } else if (beta) {
                    w = cadd(pv[e], w);
                }
                if (ok[e]) gst2(O.p, off[e], w);
            }
        }
    }
}

// One element at a time, in the order of the batched form.
template <typename R, int NACC, int WN, int NI, typename ACC>
__device__ __forceinline__ void gemm_epilogue_serial(const FbJobT<R>& J, const FbGemmArgs& g, const ACC (&acc0)[NI][WN],
                                                     const ACC (&acc1)[NI][NACC == 2 ? WN : 1], int Ncol, int n0,
                                                     int x0, int lane, int wr, int wc) {
    typedef cplx_t<R> C;
    for (int o = 0; o < J.nout; o++) {
        const FbOutT<R>& O = J.out[o];
#pragma unroll
        for (int i = 0; i < NI; i++)
#pragma unroll
            for (int j = 0; j < WN; j++) {
                const int64_t ng = n0 + wc * 16 * WN + j * 16 + (lane & 15);
                if (ng >= Ncol) continue;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int64_t xg = x0 + wr * 8 * NI + i * 8 + (sizeof(R) == 8 ? (lane >> 4) + 4 * h : 2 * (lane >> 4) + h);
                    if (xg >= g.nx) continue;
                    C v = cmul(cconv<R>(O.alpha0), cmake<R>(acc0[i][j][2 * h], acc0[i][j][2 * h + 1]));
                    if constexpr (NACC == 2)
                        v = cadd(v, cmul(cconv<R>(O.alpha1), cmake<R>(acc1[i][j][2 * h], acc1[i][j][2 * h + 1])));
                    const int64_t off = xg + g.ldo * ng;
                    if (O.diag) v = cadd(v, crs((R)g.kx[xg], cmul(cconv<R>(O.gamma), gld2(O.diag, off))));
                    if (O.scale) v = crs(gld1(O.scale, off), v);
                    if (O.poiss) {
                        const C d = crs((R)g.dt_inv, csub(gld2(O.poiss_np1, off), gld2(O.poiss_n, off)));
                        gst2(O.p, off, cadd(gld2(O.p, off), crs(gld1(O.poiss_w2, off), cadd(v, d))));
                    } else {
                        gst2(O.p, off, J.beta ? cadd(gld2(O.p, off), v) : v);
                    }
                }
            }
    }
}

// Block tile: 64 complex x (128 real rows) x 32*WN columns; 4 waves as 2 x 2,
// each 64 rows x 16*WN columns.  NACC = 2 keeps a second accumulator set,
// fed by segment 1 (launch_fb_gemm orders the segments so).  The K-stages of
// all segments form one pipeline: the global loads of stage t+1 go to
// registers while the MFMAs of stage t run, then land in the other LDS
// buffer, with one barrier per stage.
template <int NACC, int WN, int GX>
__global__ void __launch_bounds__(GB_THREADS, 2) k_fb_gemm(const FbJob* __restrict__ jobs, FbGemmArgs g) {
    constexpr int NI = GX / 16, BN = 32 * WN, LDA = 2 * GX + GB_PAD;
    constexpr int A_STAGE = GB_K * LDA, D_STAGE = BN * GB_LDK;
    const FbJob& J = jobs[blockIdx.z];
    const int Ncol = J.Ncol, K = J.K, nseg = J.nseg;
    const int n0 = blockIdx.y * BN;
    if (n0 >= Ncol) return;
    const int x0 = blockIdx.x * GX;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wr = wave & 1, wc = wave >> 1;

    extern __shared__ double gb_smem[];
    double* const As = gb_smem;                  // [2][GB_K][LDA]
    double* const Ds = gb_smem + 2 * A_STAGE;    // [2][BN][GB_LDK]

    dbl4 acc0[NI][WN], acc1[NI][NACC == 2 ? WN : 1];
#pragma unroll
    for (int i = 0; i < NI; i++)
#pragma unroll
        for (int j = 0; j < WN; j++) acc0[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

    constexpr int NA = (GB_K * GX) / GB_THREADS, ND = (GB_K * BN) / GB_THREADS;
    const int stages_per_seg = (K + GB_K - 1) / GB_K;
    const int nstages = stages_per_seg * nseg;
    const FbSeg S0 = J.seg[0];
    const FbSeg S1 = J.seg[nseg > 1 ? 1 : 0];
    double2 ra[NA], rb[NA];
    double rd[ND];
    // lane offsets of the fast-path loads, per segment
    const int xq = tid % GX, kq0 = tid / GX;
    const bool xfull = x0 + GX <= g.nx, nfull = n0 + BN <= Ncol;
    const uint32_t la0 = seg_lane_a(S0, x0 + xq, g.nx, g.lda, kq0), la1 = seg_lane_a(S1, x0 + xq, g.nx, g.lda, kq0);
    const uint32_t ld0 = (uint32_t)((tid % GB_K + S0.ldd * (tid / GB_K)) * (int64_t)sizeof(double));
    const uint32_t ld1 = (uint32_t)((tid % GB_K + S1.ldd * (tid / GB_K)) * (int64_t)sizeof(double));
    const bool plain0 = seg_plain(S0), plain1 = seg_plain(S1);
    auto fetch_seg = [&](const FbSeg& S, int k0, uint32_t la, uint32_t ldo) __attribute__((always_inline)) {
        // (not in the 128-column FP64 tiles, which have no registers left for it)
        if (WN < 4 && g.buf_ok && xfull && nfull && k0 + GB_K <= K) {
            const uint32_t row = (uint32_t)(g.lda * sizeof(double2));
            const __amdgpu_buffer_rsrc_t rs_a = ch_rsrc(S.a0);
#pragma unroll
            for (int i = 0; i < NA; i++) ra[i] = ldb_c(rs_a, la, (uint32_t)(k0 + (GB_THREADS / GX) * i) * row, double2{});
            if (S.a1) {
                const __amdgpu_buffer_rsrc_t rs_b = ch_rsrc(S.a1);
#pragma unroll
                for (int i = 0; i < NA; i++) rb[i] = ldb_c(rs_b, la, (uint32_t)(k0 + (GB_THREADS / GX) * i) * row, double2{});
            }
            const uint32_t col = (uint32_t)(S.ldd * sizeof(double));
            const __amdgpu_buffer_rsrc_t rs_d = ch_rsrc(S.D);
#pragma unroll
            for (int i = 0; i < ND; i++)
                rd[i] = ldb_r(rs_d, ldo, (uint32_t)(k0 * sizeof(double)) + (uint32_t)(n0 + (GB_THREADS / GB_K) * i) * col, double{});
            return;
        }
#pragma unroll
        for (int i = 0; i < NA; i++) {
            const int e = tid + GB_THREADS * i;
            const int xl = e % GX, kk = e / GX;
            const int64_t xg = x0 + xl;
            const bool ok = xg < g.nx && k0 + kk < K;
            const int64_t off = seg_off(S, xg, k0 + kk, g.nx, g.lda);
            ra[i] = ok ? gld2(S.a0, off) : cmk(0.0, 0.0);
        }
        if (S.a1) {
#pragma unroll
            for (int i = 0; i < NA; i++) {
                const int e = tid + GB_THREADS * i;
                const int xl = e % GX, kk = e / GX;
                const int64_t xg = x0 + xl;
                const bool ok = xg < g.nx && k0 + kk < K;
                rb[i] = ok ? gld2(S.a1, seg_off(S, xg, k0 + kk, g.nx, g.lda)) : cmk(0.0, 0.0);
            }
        }
#pragma unroll
        for (int i = 0; i < ND; i++) {
            const int e = tid + GB_THREADS * i;
            const int kk = e % GB_K, nl = e / GB_K;
            rd[i] = (k0 + kk < K && n0 + nl < Ncol) ? gld1(S.D, (int64_t)(k0 + kk) + S.ldd * (int64_t)(n0 + nl)) : 0.0;
        }
    };
    auto fetch = [&](int t) __attribute__((always_inline)) {
        const int k0 = (t % stages_per_seg) * GB_K;
        if (t < stages_per_seg) fetch_seg(S0, k0, la0, ld0);
        else fetch_seg(S1, k0, la1, ld1);
    };
    auto stash_seg = [&](const FbSeg& S, int buf, bool plain) __attribute__((always_inline)) {
        double* A = As + buf * A_STAGE;
        double* D = Ds + buf * D_STAGE;
#pragma unroll
        for (int i = 0; i < NA; i++) {
            const int e = tid + GB_THREADS * i;
            const int xl = e % GX, kk = e / GX;
            const int c = xl & 7, slot = (xl >> 3) * 16 + (c & 3) + 8 * (c >> 2);
            // (the 128-column tiles always scale: the branch costs them a spill)
            const double2 v = (WN < 4 && plain) ? ra[i] : seg_finish(S, ra[i], rb[i]);
            A[kk * LDA + slot] = v.x;
            A[kk * LDA + slot + 4] = v.y;
        }
#pragma unroll
        for (int i = 0; i < ND; i++) {
            const int e = tid + GB_THREADS * i;
            D[(e / GB_K) * GB_LDK + e % GB_K] = rd[i];
        }
    };
    auto stash = [&](int buf, int t) __attribute__((always_inline)) {
        if (t < stages_per_seg) stash_seg(S0, buf, plain0);
        else stash_seg(S1, buf, plain1);
    };
    // K-stages [tb, te) into one accumulator set; the prefetch runs across the
    // boundary between the two segments of a two-accumulator job
    auto pipe = [&](dbl4 (&acc)[NI][WN], int tb, int te) __attribute__((always_inline)) {
        for (int t = tb; t < te; t++) {
            if (t + 1 < nstages) fetch(t + 1);
            mfma_stage<WN, NI>(acc, As + (t & 1) * A_STAGE, Ds + (t & 1) * D_STAGE, lane, wr, wc);
            if (t + 1 < nstages) stash((t + 1) & 1, t + 1);
            __syncthreads();
        }
    };
    if (nstages > 0) {
        fetch(0);
        stash(0, 0);
    }
    __syncthreads();
    if constexpr (NACC == 2) {
        pipe(acc0, 0, nseg > 0 ? stages_per_seg : 0);   // a job without segments stages nothing
#pragma unroll
        for (int i = 0; i < NI; i++)
#pragma unroll
            for (int j = 0; j < WN; j++) acc1[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
        pipe(acc1, stages_per_seg, nstages);
    } else {
        pipe(acc0, 0, nstages);
    }
    gemm_epilogue<double, NACC, WN, NI>(J, g, acc0, acc1, Ncol, n0, x0, lane, wr, wc);
}

template <int NACC, int WN, int GX = GB_X>
static int32_t launch_gemm_variant(const FbJob* dj, size_t njobs, int maxN, const FbGemmArgs& g) {
    constexpr int BN = 32 * WN;
    static bool attr = false;
    if (!attr) {
        CH_HIP(hipFuncSetAttribute((const void*)k_fb_gemm<NACC, WN, GX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)gb_lds_bytes(WN, GX)));
        attr = true;
    }
    dim3 grid((unsigned)((g.nx + GX - 1) / GX), (unsigned)((maxN + BN - 1) / BN), (unsigned)njobs);
    hipLaunchKernelGGL((k_fb_gemm<NACC, WN, GX>), grid, GB_THREADS, gb_lds_bytes(WN, GX), stream(), dj, g);
    CH_LAUNCH_CHECK("k_fb_gemm");
    return ST_OK;
}

// Jobs whose segments all feed one accumulator go to `one` (with the output
// coefficients of accumulator 1 moved to 0 where needed); the others to `two`,
// ordered so that segment s feeds accumulator s.
template <typename R>
static int32_t split_jobs(const std::vector<FbJobT<R>>& jobs, std::vector<FbJobT<R>>& one,
                          std::vector<FbJobT<R>>& two) {
    for (FbJobT<R> j : jobs) {
        bool any0 = false, any1 = false;
        for (int s = 0; s < j.nseg; s++) (j.seg[s].acc ? any1 : any0) = true;
        if (any1 && !any0) {
            for (int s = 0; s < j.nseg; s++) j.seg[s].acc = 0;
            for (int o = 0; o < j.nout; o++) j.out[o].alpha0 = j.out[o].alpha1;
            any1 = false;
        }
        if (!any1) {
            one.push_back(j);
            continue;
        }
        if (j.nseg != 2) return fail(ST_EINVAL, "fb gemm job with %d segments on two accumulators", j.nseg);
        if (j.seg[0].acc == 1) std::swap(j.seg[0], j.seg[1]);
        two.push_back(j);
    }
    return ST_OK;
}

// Small grids: 64-column tiles of all the jobs together would leave CUs without a block.
template <typename J>
static bool gemm_narrow(int nx, const std::vector<J>& jobs) {
    int maxN = 0;
    for (const auto& j : jobs) maxN = j.Ncol > maxN ? j.Ncol : maxN;
    return (int64_t)((nx + GB_X - 1) / GB_X) * ((maxN + 63) / 64) * (int64_t)jobs.size() < 256;
}

// For one launch of everything: a one-accumulator job becomes a two-accumulator
// job whose second segment (if any) feeds accumulator 1 with the same output
// coefficients, so that sum_o alpha0*acc0 + alpha1*acc1 is unchanged; with one
// segment accumulator 1 stays zero.
template <typename R>
static void merge_jobs(std::vector<FbJobT<R>>& two, const std::vector<FbJobT<R>>& one) {
    for (FbJobT<R> j : one) {
        for (int o = 0; o < j.nout; o++) j.out[o].alpha1 = j.nseg == 2 ? j.out[o].alpha0 : make_double2(0.0, 0.0);
        two.push_back(j);
    }
}

// Whether every byte offset the fast-path buffer loads form (A: K rows of lda
// complex values; D: Ncol columns of ldd reals) stays below 2^31.
template <typename R>
static int gemm_offsets_fit(const std::vector<FbJobT<R>>& jobs, const FbGemmArgs& g) {
    const int64_t lim = (int64_t)1 << 31;
    for (const auto& j : jobs)
        for (int s = 0; s < j.nseg; s++) {
            if (((int64_t)j.K * g.lda + g.nx) * (int64_t)sizeof(cplx_t<R>) >= lim) return 0;
            if (((int64_t)j.K + j.seg[s].ldd * (int64_t)j.Ncol) * (int64_t)sizeof(R) >= lim) return 0;
        }
    return 1;
}

// Tile of a launch: the first of 128-, 64- and 32-column tiles of 64-complex x
// and 32-column tiles of 32- and 16-complex x that makes at least
// GEMM_MIN_BLOCKS blocks (two per CU), else the last.  Larger tiles re-read
// less of A and D; smaller ones put waves on every SIMD when the grid is small.
constexpr int64_t GEMM_MIN_BLOCKS = 512;
enum { GT_W4 = 0, GT_W2, GT_W1, GT_H, GT_Q };
static int gemm_tile(int nx, int maxN, size_t njobs, bool has_w4) {
    const int64_t xb = (nx + GB_X - 1) / GB_X, nj = (int64_t)njobs, c32 = (maxN + 31) / 32;
    const int64_t nb[GT_Q] = {xb * ((maxN + 127) / 128) * nj, xb * ((maxN + 63) / 64) * nj, xb * c32 * nj,
                              ((nx + GB_X / 2 - 1) / (GB_X / 2)) * c32 * nj};
    for (int t = has_w4 ? GT_W4 : GT_W2; t < GT_Q; t++)
        if (nb[t] >= GEMM_MIN_BLOCKS) return t;
    return GT_Q;
}

template <int NACC, int WN, int GX = GB_X>
static int32_t launch_gemm_variant_f32(const FbJobT<float>* dj, size_t njobs, int maxN, const FbGemmArgs& g);
template <int NACC, int WN, int GX = GB_X>
static int32_t launch_gemm_variant_bf16(const FbJobT<float>* dj, size_t njobs, int maxN, const FbGemmArgs& g);

// Form of the mixed-precision GEMMs: split-BF16 (six products, the default)
// or FP32 MFMA; process-wide, set by chimera_set_mixed_gemm.
static std::atomic<int> g_mixed_gemm{MIXED_GEMM_BF16X6};

// One launch of a job table in tile shape `tile` (GT_*) with NACC
// accumulators; EUNSUP for a shape the form has no kernel for.
static int32_t gemm_tile_launch(const FbJob* dj, size_t nj, int maxN, const FbGemmArgs& g, int nacc, int tile) {
    if (nacc == 1) {
        switch (tile) {
            case GT_W4: return launch_gemm_variant<1, 4>(dj, nj, maxN, g);
            case GT_W2: return launch_gemm_variant<1, 2>(dj, nj, maxN, g);
            case GT_W1: return launch_gemm_variant<1, 1>(dj, nj, maxN, g);
            case GT_H: return launch_gemm_variant<1, 1, GB_X / 2>(dj, nj, maxN, g);
            default: return launch_gemm_variant<1, 1, GB_X / 4>(dj, nj, maxN, g);
        }
    }
    switch (tile) {
        case GT_W4: return ST_EUNSUP;
        case GT_W2: return launch_gemm_variant<2, 2>(dj, nj, maxN, g);
        case GT_W1: return launch_gemm_variant<2, 1>(dj, nj, maxN, g);
        case GT_H: return launch_gemm_variant<2, 1, GB_X / 2>(dj, nj, maxN, g);
        default: return launch_gemm_variant<2, 1, GB_X / 4>(dj, nj, maxN, g);
    }
}

static int32_t gemm_tile_launch(const FbJobT<float>* dj, size_t nj, int maxN, const FbGemmArgs& g, int nacc,
                                int tile) {
    if (g_mixed_gemm.load() == MIXED_GEMM_BF16X6) {
        if (nacc == 1) {
            switch (tile) {
                case GT_W4: return launch_gemm_variant_bf16<1, 4>(dj, nj, maxN, g);
                case GT_W2: return launch_gemm_variant_bf16<1, 2>(dj, nj, maxN, g);
                case GT_W1: return launch_gemm_variant_bf16<1, 1>(dj, nj, maxN, g);
                case GT_H: return launch_gemm_variant_bf16<1, 1, GB_X / 2>(dj, nj, maxN, g);
                default: return launch_gemm_variant_bf16<1, 1, GB_X / 4>(dj, nj, maxN, g);
            }
        }
        switch (tile) {
            case GT_W4: return ST_EUNSUP;          // two accumulators of 128 columns: no registers left
            case GT_W2: return launch_gemm_variant_bf16<2, 2>(dj, nj, maxN, g);
            case GT_W1: return launch_gemm_variant_bf16<2, 1>(dj, nj, maxN, g);
            case GT_H: return launch_gemm_variant_bf16<2, 1, GB_X / 2>(dj, nj, maxN, g);
            default: return launch_gemm_variant_bf16<2, 1, GB_X / 4>(dj, nj, maxN, g);
        }
    }
    if (nacc == 1) {
        switch (tile) {
            case GT_W4: return launch_gemm_variant_f32<1, 4>(dj, nj, maxN, g);
            case GT_W2: return launch_gemm_variant_f32<1, 2>(dj, nj, maxN, g);
            case GT_W1: return launch_gemm_variant_f32<1, 1>(dj, nj, maxN, g);
            case GT_H: return launch_gemm_variant_f32<1, 1, GB_X / 2>(dj, nj, maxN, g);
            default: return launch_gemm_variant_f32<1, 1, GB_X / 4>(dj, nj, maxN, g);
        }
    }
    switch (tile) {
        case GT_W4: return launch_gemm_variant_f32<2, 4>(dj, nj, maxN, g);
        case GT_W2: return launch_gemm_variant_f32<2, 2>(dj, nj, maxN, g);
        case GT_W1: return launch_gemm_variant_f32<2, 1>(dj, nj, maxN, g);
        case GT_H: return launch_gemm_variant_f32<2, 1, GB_X / 2>(dj, nj, maxN, g);
        default: return launch_gemm_variant_f32<2, 1, GB_X / 4>(dj, nj, maxN, g);
    }
}

// ---------------------------------------------------------------- tile autotuner
//
// With autotuning on (chimera_set_autotune), the first launch of a GEMM shape
// -- (element type, GEMM form, accumulators, nx, columns, jobs, K, segments,
// buffer-load path) -- times every tile shape the form has a kernel for and
// keeps the fastest for that shape; later launches of the shape use it.  The
// timed launches run the real job table with every output redirected to a
// scratch buffer (the operands are only read), three launches per shape after
// one untimed, HIP events on the library stream; the shape with the least
// mean time wins.  Tuned shapes can be kept in a text file
// (chimera_set_autotune's path: one "key... tile mean_ms" line per shape,
// read once and appended to).  Under stream capture, or when the scratch
// cannot grow, a shape not yet tuned takes the block-count rule (gemm_tile).
// Every tile shape sums each output over K in the same order, so the choice
// changes the time of a launch, not its result.
struct TuneKey {
    int v[9];
    bool operator<(const TuneKey& o) const { return std::lexicographical_compare(v, v + 9, o.v, o.v + 9); }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, int> g_tuned;
static std::atomic<int> g_autotune{0};
static std::string g_tune_path;
static bool g_tune_loaded = false;

static void tune_load_locked() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    if (g_tune_path.empty()) return;
    FILE* f = fopen(g_tune_path.c_str(), "r");
    if (!f) return;
    TuneKey k;
    int tile;
    double ms;
    while (fscanf(f, "%d %d %d %d %d %d %d %d %d %d %lf", &k.v[0], &k.v[1], &k.v[2], &k.v[3], &k.v[4], &k.v[5],
                  &k.v[6], &k.v[7], &k.v[8], &tile, &ms) == 11)
        if (tile >= GT_W4 && tile <= GT_Q) g_tuned[k] = tile;
    fclose(f);
}

template <typename J>
static int gemm_choose_tile(const J* dj, const std::vector<J>& js, int maxN, const FbGemmArgs& g, int nacc,
                            int rbytes) {
    const size_t nj = js.size();
    const int rule = gemm_tile(g.nx, maxN, nj, rbytes == 4 || nacc == 1);
    if (g_autotune.load() == 0) return rule;
    int maxK = 0, nseg = 0;
    for (const auto& j : js) {
        maxK = j.K > maxK ? j.K : maxK;
        nseg = j.nseg > nseg ? j.nseg : nseg;
    }
    const int form = rbytes == 4 ? g_mixed_gemm.load() : 2;
    const TuneKey key = {{rbytes, form, nacc, g.nx, maxN, (int)nj, maxK, nseg, g.buf_ok}};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load_locked();
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) return it->second;
    }
    if (capturing()) return rule;
    // the job table with every output in one scratch buffer of ldo * maxN elements
    std::vector<J> tj(js);
    void* sink = scratch(20, (size_t)g.ldo * maxN * 2 * rbytes);
    if (!sink) return rule;
    for (auto& j : tj)
        for (int o = 0; o < j.nout; o++) j.out[o].p = static_cast<decltype(j.out[o].p)>(sink);
    const J* tdj = static_cast<const J*>(device_table(tj.data(), tj.size() * sizeof(J)));
    if (!tdj) return rule;
    (void)dj;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return rule;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return rule;
    }
    int best = rule;
    float best_ms = 1e30f;
    for (int t = GT_W4; t <= GT_Q; t++) {
        if (gemm_tile_launch(tdj, nj, maxN, g, nacc, t) != ST_OK) continue;     // warm-up (or no such kernel)
        (void)hipEventRecord(e0, stream());
        bool ok = true;
        for (int r = 0; r < 3 && ok; r++) ok = gemm_tile_launch(tdj, nj, maxN, g, nacc, t) == ST_OK;
        (void)hipEventRecord(e1, stream());
        if (!ok || hipEventSynchronize(e1) != hipSuccess) continue;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) continue;
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key] = best;
    if (!g_tune_path.empty()) {
        if (FILE* f = fopen(g_tune_path.c_str(), "a")) {
            for (int i = 0; i < 9; i++) fprintf(f, "%d ", key.v[i]);
            fprintf(f, "%d %.6f\n", best, best_ms / 3.0);
            fclose(f);
        }
    }
    return best;
}

template <>
int32_t launch_fb_gemm<double>(const std::vector<FbJob>& jobs, const FbGemmArgs& g_in, int family) {
    if (jobs.empty()) return ST_OK;
    KTimer kt(family);
    FbGemmArgs g = g_in;
    g.buf_ok = gemm_offsets_fit(jobs, g);
    std::vector<FbJob> one, two;
    CH_TRY(split_jobs<double>(jobs, one, two));
    if (!two.empty() && !one.empty() && gemm_narrow(g.nx, jobs)) {
        merge_jobs<double>(two, one);   // small grids: one launch instead of two
        one.clear();
    }
    for (int pass = 0; pass < 2; pass++) {
        const std::vector<FbJob>& js = pass == 0 ? one : two;
        if (js.empty()) continue;
        int maxN = 0;
        for (const auto& j : js) maxN = j.Ncol > maxN ? j.Ncol : maxN;
        const FbJob* dj = static_cast<const FbJob*>(device_table(js.data(), js.size() * sizeof(FbJob)));
        if (!dj) return ST_ENOMEM;
        const int tile = gemm_choose_tile<FbJob>(dj, js, maxN, g, pass + 1, 8);
        CH_TRY(gemm_tile_launch(dj, js.size(), maxN, g, pass + 1, tile));
    }
    return ST_OK;
}

// ---------------------------------------------------------------- FP32 (mixed precision)
//
// Same job structure on complex64 data with v_mfma_f32_16x16x4f32.  Lane
// layout on gfx950 (checked on the device): lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and holds D[4*(l/16) + i][l%16], i = 0..3.  A rows stay in
// memory order (re, im interleaved), so a lane's items (0,1) and (2,3) are the
// (re, im) of x = 2*(l/16) and 2*(l/16) + 1 within the 8-complex MFMA tile.
// Accumulators take half the registers of FP64, so both accumulator sets fit
// with 64x64 wave tiles.
typedef float flt4 __attribute__((ext_vector_type(4)));
// K-stage depth of the FP32 kernel: 16 like the FP64 kernel, and 32 for the
// 32-column tiles of small grids, whose launches are too few blocks to hide
// the global-load latency of a stage and gain from half as many stages.
constexpr int gbf_k(int nacc, int wn) { return wn == 1 ? 32 : 16; }

template <int WN, int GBF_K, int NI>
__device__ __forceinline__ void mfma_stage_f32(flt4 (&acc)[NI][WN], const float* __restrict__ As,
                                               const float* __restrict__ Ds, int lane, int wr, int wc) {
    constexpr int LDA = 32 * NI + GB_PAD, LDD = 32 * WN + GB_PAD;
#pragma unroll
    for (int kq = 0; kq < GBF_K / 4; kq++) {
        const int kk = kq * 4 + (lane >> 4);
        float a[NI], b[WN];
#pragma unroll
        for (int i = 0; i < NI; i++) a[i] = As[kk * LDA + wr * 16 * NI + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < WN; j++) b[j] = Ds[kk * LDD + wc * 16 * WN + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < NI; i++)
#pragma unroll
            for (int j = 0; j < WN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
}

constexpr size_t gb_lds_bytes_f32(int WN, int GBF_K, int GX) {
    return 2 * (size_t)GBF_K * ((2 * GX + GB_PAD) + (32 * WN + GB_PAD)) * sizeof(float);
}

template <int NACC, int WN, int GX>
__global__ void __launch_bounds__(GB_THREADS, 2) k_fb_gemm_f32(const FbJobT<float>* __restrict__ jobs, FbGemmArgs g) {
    typedef FbSegT<float> Seg;
    constexpr int GBF_K = gbf_k(NACC, WN);
    constexpr int NI = GX / 16, BN = 32 * WN, LDA = 2 * GX + GB_PAD, LDD = BN + GB_PAD;
    constexpr int A_STAGE = GBF_K * LDA, D_STAGE = GBF_K * LDD;
    const FbJobT<float>& J = jobs[blockIdx.z];
    const int Ncol = J.Ncol, K = J.K, nseg = J.nseg;
    const int n0 = blockIdx.y * BN;
    if (n0 >= Ncol) return;
    const int x0 = blockIdx.x * GX;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wr = wave & 1, wc = wave >> 1;

    extern __shared__ __align__(16) unsigned char gbf_smem[];
    float* const As = reinterpret_cast<float*>(gbf_smem);  // [2][GBF_K][LDA]
    float* const Ds = As + 2 * A_STAGE;                       // [2][GBF_K][LDD]

    flt4 acc0[NI][WN], acc1[NI][NACC == 2 ? WN : 1];
#pragma unroll
    for (int i = 0; i < NI; i++)
#pragma unroll
        for (int j = 0; j < WN; j++) acc0[i][j] = flt4{0.f, 0.f, 0.f, 0.f};

    constexpr int NA = (GBF_K * GX) / GB_THREADS, ND = (GBF_K * BN) / GB_THREADS;
    const int stages_per_seg = (K + GBF_K - 1) / GBF_K;
    const int nstages = stages_per_seg * nseg;
    const Seg S0 = J.seg[0];
    const Seg S1 = J.seg[nseg > 1 ? 1 : 0];
    float2 ra[NA], rb[NA];
    float rd[ND];
    // lane offsets of the fast-path loads, per segment
    const int xq = tid % GX, kq0 = tid / GX;
    const bool xfull = x0 + GX <= g.nx, nfull = n0 + BN <= Ncol;
    const uint32_t la0 = seg_lane_a(S0, x0 + xq, g.nx, g.lda, kq0), la1 = seg_lane_a(S1, x0 + xq, g.nx, g.lda, kq0);
    const uint32_t ld0 = (uint32_t)((tid % GBF_K + S0.ldd * (tid / GBF_K)) * (int64_t)sizeof(float));
    const uint32_t ld1 = (uint32_t)((tid % GBF_K + S1.ldd * (tid / GBF_K)) * (int64_t)sizeof(float));
    const bool plain0 = seg_plain(S0), plain1 = seg_plain(S1);
    auto fetch_seg = [&](const Seg& S, int k0, uint32_t la, uint32_t ldo) __attribute__((always_inline)) {
        if (g.buf_ok && xfull && nfull && k0 + GBF_K <= K) {
            const uint32_t row = (uint32_t)(g.lda * sizeof(float2));
            const __amdgpu_buffer_rsrc_t rs_a = ch_rsrc(S.a0);
#pragma unroll
            for (int i = 0; i < NA; i++) ra[i] = ldb_c(rs_a, la, (uint32_t)(k0 + (GB_THREADS / GX) * i) * row, float2{});
            if (S.a1) {
                const __amdgpu_buffer_rsrc_t rs_b = ch_rsrc(S.a1);
#pragma unroll
                for (int i = 0; i < NA; i++) rb[i] = ldb_c(rs_b, la, (uint32_t)(k0 + (GB_THREADS / GX) * i) * row, float2{});
            }
            const uint32_t col = (uint32_t)(S.ldd * sizeof(float));
            const __amdgpu_buffer_rsrc_t rs_d = ch_rsrc(S.D);
#pragma unroll
            for (int i = 0; i < ND; i++)
                rd[i] = ldb_r(rs_d, ldo, (uint32_t)(k0 * sizeof(float)) + (uint32_t)(n0 + (GB_THREADS / GBF_K) * i) * col, float{});
            return;
        }
#pragma unroll
        for (int i = 0; i < NA; i++) {
            const int e = tid + GB_THREADS * i;
            const int xl = e % GX, kk = e / GX;
            const int64_t xg = x0 + xl;
            const bool ok = xg < g.nx && k0 + kk < K;
            ra[i] = ok ? gld2(S.a0, seg_off(S, xg, k0 + kk, g.nx, g.lda)) : cmkf(0.f, 0.f);
        }
        if (S.a1) {
#pragma unroll
            for (int i = 0; i < NA; i++) {
                const int e = tid + GB_THREADS * i;
                const int xl = e % GX, kk = e / GX;
                const int64_t xg = x0 + xl;
                const bool ok = xg < g.nx && k0 + kk < K;
                rb[i] = ok ? gld2(S.a1, seg_off(S, xg, k0 + kk, g.nx, g.lda)) : cmkf(0.f, 0.f);
            }
        }
#pragma unroll
        for (int i = 0; i < ND; i++) {
            const int e = tid + GB_THREADS * i;
            const int kk = e % GBF_K, nl = e / GBF_K;
            rd[i] = (k0 + kk < K && n0 + nl < Ncol) ? gld1(S.D, (int64_t)(k0 + kk) + S.ldd * (int64_t)(n0 + nl)) : 0.f;
        }
    };
    auto fetch = [&](int t) __attribute__((always_inline)) {
        const int k0 = (t % stages_per_seg) * GBF_K;
        if (t < stages_per_seg) fetch_seg(S0, k0, la0, ld0);
        else fetch_seg(S1, k0, la1, ld1);
    };
    auto stash_seg = [&](const Seg& S, int buf, bool plain) __attribute__((always_inline)) {
        float* A = As + buf * A_STAGE;
        float* D = Ds + buf * D_STAGE;
#pragma unroll
        for (int i = 0; i < NA; i++) {
            const int e = tid + GB_THREADS * i;
            const int xl = e % GX, kk = e / GX;
            const float2 v = plain ? ra[i] : seg_finish(S, ra[i], rb[i]);
            *reinterpret_cast<float2*>(A + kk * LDA + 2 * xl) = v;
        }
#pragma unroll
        for (int i = 0; i < ND; i++) {
            const int e = tid + GB_THREADS * i;
            D[(e % GBF_K) * LDD + e / GBF_K] = rd[i];
        }
    };
    auto stash = [&](int buf, int t) __attribute__((always_inline)) {
        if (t < stages_per_seg) stash_seg(S0, buf, plain0);
        else stash_seg(S1, buf, plain1);
    };
    auto pipe = [&](flt4 (&acc)[NI][WN], int tb, int te) __attribute__((always_inline)) {
        for (int t = tb; t < te; t++) {
            if (t + 1 < nstages) fetch(t + 1);
            mfma_stage_f32<WN, GBF_K, NI>(acc, As + (t & 1) * A_STAGE, Ds + (t & 1) * D_STAGE, lane, wr, wc);
            if (t + 1 < nstages) stash((t + 1) & 1, t + 1);
            __syncthreads();
        }
    };
    if (nstages > 0) {
        fetch(0);
        stash(0, 0);
    }
    __syncthreads();
    if constexpr (NACC == 2) {
        pipe(acc0, 0, nseg > 0 ? stages_per_seg : 0);   // a job without segments stages nothing
#pragma unroll
        for (int i = 0; i < NI; i++)
#pragma unroll
            for (int j = 0; j < WN; j++) acc1[i][j] = flt4{0.f, 0.f, 0.f, 0.f};
        pipe(acc1, stages_per_seg, nstages);
    } else {
        pipe(acc0, 0, nstages);
    }
    gemm_epilogue<float, NACC, WN, NI>(J, g, acc0, acc1, Ncol, n0, x0, lane, wr, wc);
}

template <int NACC, int WN, int GX>
static int32_t launch_gemm_variant_f32(const FbJobT<float>* dj, size_t njobs, int maxN, const FbGemmArgs& g) {
    constexpr int BN = 32 * WN;
    constexpr size_t lds = gb_lds_bytes_f32(WN, gbf_k(NACC, WN), GX);
    static bool attr = false;
    if (!attr) {
        CH_HIP(hipFuncSetAttribute((const void*)k_fb_gemm_f32<NACC, WN, GX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
        attr = true;
    }
    dim3 grid((unsigned)((g.nx + GX - 1) / GX), (unsigned)((maxN + BN - 1) / BN), (unsigned)njobs);
    hipLaunchKernelGGL((k_fb_gemm_f32<NACC, WN, GX>), grid, GB_THREADS, lds, stream(), dj, g);
    CH_LAUNCH_CHECK("k_fb_gemm_f32");
    return ST_OK;
}

// ---------------------------------------------------------------- split-BF16 (mixed precision)
//
// The FP32 GEMM of the mixed path on the BF16 matrix cores: every FP32 value
// v of A (fields) and D (matrices) is split exactly into three BF16 pieces,
// v = h + m + l (h = RNE(v) to 8 significant bits, m = RNE(v - h), l = v - h - m;
// the differences are exact in FP32), and
//   A.D ~= Ah.Dh + Ah.Dm + Am.Dh + Ah.Dl + Am.Dm + Al.Dh
// in six v_mfma_f32_16x16x32_bf16 products with FP32 accumulation.  The
// dropped products (Am.Dl, Al.Dm, Al.Dl) are below 2^-24 of A.D, so the
// result has the accuracy of the FP32 kernel; three products (dropping the
// 2^-16 terms) would leave ~1e-5 relative error per transform.
//
// Lane layout of v_mfma_f32_16x16x32_bf16 on gfx950 (checked on the device):
// lane l supplies A[l%16][8*(l/16) + j] and B[8*(l/16) + j][l%16], j = 0..7,
// and holds D[4*(l/16) + i][l%16], i = 0..3 (the output layout of the FP32
// kernel, so both share the epilogue).  A rows stay in the FP32 kernel's
// order (re, im of x interleaved).  K is staged 32 at a time; the pieces are
// kept in LDS with K contiguous per row (A) and per column (D), so that each
// operand is one 16-byte LDS read, rows 80 bytes apart (16-lane groups of a
// 16-byte read then hit disjoint banks).  One LDS buffer: the next stage is
// loaded into registers while the MFMAs of this one run.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int GBB_K = 32, GBB_RS = GBB_K + 8;          // K per stage; LDS row stride in BF16

// v = h + m + l exactly (finite v of normal magnitude); pieces as BF16 bit
// patterns.  Each piece is the round-to-nearest-even BF16 of what is left
// (gfx950's v_cvt_pk_bf16_f32); the remainders are exact in FP32, since a
// piece keeps the leading 8 bits of its remainder.
__device__ __forceinline__ void split3_bf16(float v, uint16_t& h, uint16_t& m, uint16_t& l) {
    const __bf16 bh = (__bf16)v;
    const float r1 = v - (float)bh;
    const __bf16 bm = (__bf16)r1;
    const float r2 = r1 - (float)bm;
    h = __builtin_bit_cast(uint16_t, bh);
    m = __builtin_bit_cast(uint16_t, bm);
    l = __builtin_bit_cast(uint16_t, (__bf16)r2);
}

// n consecutive BF16 bit patterns (n = 2, 4 or 8; dst aligned to 2n bytes)
// as one LDS store
template <int n>
__device__ __forceinline__ void lds_store_bf16(uint16_t* dst, const uint16_t (&v)[n]) {
    static_assert(n == 2 || n == 4 || n == 8, "2, 4 or 8 pieces");
    uint32_t w[n / 2];
#pragma unroll
    for (int i = 0; i < n / 2; i++) w[i] = (uint32_t)v[2 * i] | ((uint32_t)v[2 * i + 1] << 16);
    if constexpr (n == 8) *reinterpret_cast<uint4*>(dst) = make_uint4(w[0], w[1], w[2], w[3]);
    else if constexpr (n == 4) *reinterpret_cast<uint2*>(dst) = make_uint2(w[0], w[1]);
    else *reinterpret_cast<uint32_t*>(dst) = w[0];
}

constexpr size_t gb_lds_bytes_bf16(int WN, int GX) {
    return 3 * (size_t)(2 * GX + 32 * WN) * GBB_RS * sizeof(uint16_t);
}

template <int NACC, int WN, int GX>
__global__ void __launch_bounds__(GB_THREADS, 2) k_fb_gemm_bf16(const FbJobT<float>* __restrict__ jobs, FbGemmArgs g) {
    typedef FbSegT<float> Seg;
    constexpr int NI = GX / 16, BN = 32 * WN, ROWS = 2 * GX;
    constexpr int KA = GBB_K * GX / GB_THREADS;          // consecutive k per thread in the A fetch
    constexpr int DC = (BN * 4 + GB_THREADS - 1) / GB_THREADS;   // D columns per thread (8 k each)
    const FbJobT<float>& J = jobs[blockIdx.z];
    const int Ncol = J.Ncol, K = J.K, nseg = J.nseg;
    const int n0 = blockIdx.y * BN;
    if (n0 >= Ncol) return;
    const int x0 = blockIdx.x * GX;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wr = wave & 1, wc = wave >> 1;

    extern __shared__ __align__(16) unsigned char gbb_smem[];
    uint16_t* const As = reinterpret_cast<uint16_t*>(gbb_smem);     // [3][ROWS][GBB_RS]
    uint16_t* const Ds = As + 3 * ROWS * GBB_RS;                     // [3][BN][GBB_RS]

    flt4 acc0[NI][WN], acc1[NI][NACC == 2 ? WN : 1];
#pragma unroll
    for (int i = 0; i < NI; i++)
#pragma unroll
        for (int j = 0; j < WN; j++) acc0[i][j] = flt4{0.f, 0.f, 0.f, 0.f};

    const int stages_per_seg = (K + GBB_K - 1) / GBB_K;
    const int nstages = stages_per_seg * nseg;
    const Seg S0 = J.seg[0];
    const Seg S1 = J.seg[nseg > 1 ? 1 : 0];
    // A fetch: x = x0 + tid % GX, k = k0 + (tid / GX) * KA + q, q < KA
    const int xl = tid % GX, kb = (tid / GX) * KA;
    // D fetch: column n0 + (tid / 4) + (GB_THREADS / 4) * c, k = k0 + 8 * (tid % 4) + q, q < 8
    const int dk = 8 * (tid % 4), dn = tid / 4;
    float2 ra[KA], rb[KA];
    float rd[DC][8];
    const bool xfull = x0 + GX <= g.nx, nfull = n0 + BN <= Ncol;
    const uint32_t la0 = seg_lane_a(S0, x0 + xl, g.nx, g.lda, kb), la1 = seg_lane_a(S1, x0 + xl, g.nx, g.lda, kb);
    const bool plain0 = seg_plain(S0), plain1 = seg_plain(S1);
    auto fetch_seg = [&](const Seg& S, int k0, uint32_t la) __attribute__((always_inline)) {
        if (g.buf_ok && xfull && nfull && k0 + GBB_K <= K) {
            const uint32_t row = (uint32_t)(g.lda * sizeof(float2));
            const __amdgpu_buffer_rsrc_t rs_a = ch_rsrc(S.a0);
#pragma unroll
            for (int q = 0; q < KA; q++) ra[q] = ldb_c(rs_a, la, (uint32_t)(k0 + q) * row, float2{});
            if (S.a1) {
                const __amdgpu_buffer_rsrc_t rs_b = ch_rsrc(S.a1);
#pragma unroll
                for (int q = 0; q < KA; q++) rb[q] = ldb_c(rs_b, la, (uint32_t)(k0 + q) * row, float2{});
            }
            const __amdgpu_buffer_rsrc_t rs_d = ch_rsrc(S.D);
#pragma unroll
            for (int c = 0; c < DC; c++) {
                const int nl = dn + (GB_THREADS / 4) * c;
                const uint32_t off = (uint32_t)((k0 + dk + S.ldd * (int64_t)(n0 + (nl < BN ? nl : 0))) * sizeof(float));
#pragma unroll
                for (int q = 0; q < 8; q++) rd[c][q] = ldb_r(rs_d, off, (uint32_t)(q * sizeof(float)), float{});
            }
            return;
        }
        const int64_t xg = x0 + xl;
#pragma unroll
        for (int q = 0; q < KA; q++) {
            const bool ok = xg < g.nx && k0 + kb + q < K;
            ra[q] = ok ? gld2(S.a0, seg_off(S, xg, k0 + kb + q, g.nx, g.lda)) : cmkf(0.f, 0.f);
        }
        if (S.a1) {
#pragma unroll
            for (int q = 0; q < KA; q++) {
                const bool ok = xg < g.nx && k0 + kb + q < K;
                rb[q] = ok ? gld2(S.a1, seg_off(S, xg, k0 + kb + q, g.nx, g.lda)) : cmkf(0.f, 0.f);
            }
        }
#pragma unroll
        for (int c = 0; c < DC; c++) {
            const int nl = dn + (GB_THREADS / 4) * c;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int kk = k0 + dk + q;
                rd[c][q] = (nl < BN && kk < K && n0 + nl < Ncol) ? gld1(S.D, (int64_t)kk + S.ldd * (int64_t)(n0 + nl))
                                                                  : 0.f;
            }
        }
    };
    auto fetch = [&](int t) __attribute__((always_inline)) {
        const int k0 = (t % stages_per_seg) * GBB_K;
        if (t < stages_per_seg) fetch_seg(S0, k0, la0);
        else fetch_seg(S1, k0, la1);
    };
    // split and store: A rows 2*xl (re) and 2*xl + 1 (im), k = kb..kb+KA-1; D columns, k = dk..dk+7
    auto stash_seg = [&](const Seg& S, bool plain) __attribute__((always_inline)) {
        uint16_t hr[KA], mr[KA], lr[KA], hi[KA], mi[KA], li[KA];
#pragma unroll
        for (int q = 0; q < KA; q++) {
            const float2 v = plain ? ra[q] : seg_finish(S, ra[q], rb[q]);
            split3_bf16(v.x, hr[q], mr[q], lr[q]);
            split3_bf16(v.y, hi[q], mi[q], li[q]);
        }
        // (rows 80 bytes apart and kb a multiple of KA: every store is aligned to its size)
        uint16_t* const a_re = As + (2 * xl) * GBB_RS + kb;
        uint16_t* const a_im = a_re + GBB_RS;
        constexpr int PA = ROWS * GBB_RS;
        lds_store_bf16<KA>(a_re, hr);
        lds_store_bf16<KA>(a_re + PA, mr);
        lds_store_bf16<KA>(a_re + 2 * PA, lr);
        lds_store_bf16<KA>(a_im, hi);
        lds_store_bf16<KA>(a_im + PA, mi);
        lds_store_bf16<KA>(a_im + 2 * PA, li);
        constexpr int PD = BN * GBB_RS;
#pragma unroll
        for (int c = 0; c < DC; c++) {
            const int nl = dn + (GB_THREADS / 4) * c;
            if (nl >= BN) continue;
            uint16_t* const d = Ds + nl * GBB_RS + dk;
            uint16_t h[8], m[8], l[8];
#pragma unroll
            for (int q = 0; q < 8; q++) split3_bf16(rd[c][q], h[q], m[q], l[q]);
            lds_store_bf16<8>(d, h);
            lds_store_bf16<8>(d + PD, m);
            lds_store_bf16<8>(d + 2 * PD, l);
        }
    };
    auto stash = [&](int t) __attribute__((always_inline)) {
        if (t < stages_per_seg) stash_seg(S0, plain0);
        else stash_seg(S1, plain1);
    };
    auto mfma_stage = [&](flt4 (&acc)[NI][WN]) __attribute__((always_inline)) {
        constexpr int PA = ROWS * GBB_RS, PD = BN * GBB_RS;
        const int ko = 8 * (lane >> 4);
        bf16x8 ah[NI], am[NI], al[NI];
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const uint16_t* a = As + (wr * GX + 16 * i + (lane & 15)) * GBB_RS + ko;
            ah[i] = *reinterpret_cast<const bf16x8*>(a);
            am[i] = *reinterpret_cast<const bf16x8*>(a + PA);
            al[i] = *reinterpret_cast<const bf16x8*>(a + 2 * PA);
        }
#pragma unroll
        for (int j = 0; j < WN; j++) {
            const uint16_t* d = Ds + (wc * 16 * WN + 16 * j + (lane & 15)) * GBB_RS + ko;
            const bf16x8 bh = *reinterpret_cast<const bf16x8*>(d);
            const bf16x8 bm = *reinterpret_cast<const bf16x8*>(d + PD);
            const bf16x8 bl = *reinterpret_cast<const bf16x8*>(d + 2 * PD);
#pragma unroll
            for (int i = 0; i < NI; i++) {
                flt4 c = acc[i][j];
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[i], bm, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[i], bh, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bm, c, 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, c, 0, 0, 0);
            }
        }
    };
    auto pipe = [&](flt4 (&acc)[NI][WN], int tb, int te) __attribute__((always_inline)) {
        for (int t = tb; t < te; t++) {
            if (t + 1 < nstages) fetch(t + 1);
            mfma_stage(acc);
            __syncthreads();
            if (t + 1 < nstages) stash(t + 1);
            __syncthreads();
        }
    };
    if (nstages > 0) {
        fetch(0);
        stash(0);
    }
    __syncthreads();
    if constexpr (NACC == 2) {
        pipe(acc0, 0, nseg > 0 ? stages_per_seg : 0);   // a job without segments stages nothing
#pragma unroll
        for (int i = 0; i < NI; i++)
#pragma unroll
            for (int j = 0; j < WN; j++) acc1[i][j] = flt4{0.f, 0.f, 0.f, 0.f};
        pipe(acc1, stages_per_seg, nstages);
    } else {
        pipe(acc0, 0, nstages);
    }
    gemm_epilogue<float, NACC, WN, NI>(J, g, acc0, acc1, Ncol, n0, x0, lane, wr, wc);
}

template <int NACC, int WN, int GX>
static int32_t launch_gemm_variant_bf16(const FbJobT<float>* dj, size_t njobs, int maxN, const FbGemmArgs& g) {
    constexpr int BN = 32 * WN;
    constexpr size_t lds = gb_lds_bytes_bf16(WN, GX);
    static bool attr = false;
    if (!attr) {
        CH_HIP(hipFuncSetAttribute((const void*)k_fb_gemm_bf16<NACC, WN, GX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
        attr = true;
    }
    dim3 grid((unsigned)((g.nx + GX - 1) / GX), (unsigned)((maxN + BN - 1) / BN), (unsigned)njobs);
    hipLaunchKernelGGL((k_fb_gemm_bf16<NACC, WN, GX>), grid, GB_THREADS, lds, stream(), dj, g);
    CH_LAUNCH_CHECK("k_fb_gemm_bf16");
    return ST_OK;
}

template <>
int32_t launch_fb_gemm<float>(const std::vector<FbJobT<float>>& jobs, const FbGemmArgs& g_in, int family) {
    if (jobs.empty()) return ST_OK;
    KTimer kt(family);
    FbGemmArgs g = g_in;
    g.buf_ok = gemm_offsets_fit(jobs, g);
    std::vector<FbJobT<float>> one, two;
    CH_TRY(split_jobs<float>(jobs, one, two));
    if (!two.empty() && !one.empty() && gemm_narrow(g.nx, jobs)) {
        merge_jobs<float>(two, one);    // small grids: one launch instead of two
        one.clear();
    }
    for (int pass = 0; pass < 2; pass++) {
        const std::vector<FbJobT<float>>& js = pass == 0 ? one : two;
        if (js.empty()) continue;
        int maxN = 0;
        for (const auto& j : js) maxN = j.Ncol > maxN ? j.Ncol : maxN;
        const FbJobT<float>* dj =
            static_cast<const FbJobT<float>*>(device_table(js.data(), js.size() * sizeof(FbJobT<float>)));
        if (!dj) return ST_ENOMEM;
        int tile = gemm_choose_tile<FbJobT<float>>(dj, js, maxN, g, pass + 1, 4);
        if (tile == GT_W4 && pass == 1 && g_mixed_gemm.load() == MIXED_GEMM_BF16X6) tile = GT_W2;
        CH_TRY(gemm_tile_launch(dj, js.size(), maxN, g, pass + 1, tile));
    }
    return ST_OK;
}

// ================================================================== job builders

template <typename R>
static FbSegT<R> seg(const cplx_t<R>* a0, double2 c0, const cplx_t<R>* a1, double2 c1, const R* D, int64_t ldd,
                     int refl, int acc) {
    FbSegT<R> s;
    s.a0 = a0;
    s.a1 = a1;
    s.c0 = c0;
    s.c1 = c1;
    s.D = D;
    s.ldd = ldd;
    s.refl = refl;
    s.acc = acc;
    return s;
}

template <typename R>
static FbOutT<R> outp(cplx_t<R>* p, double2 al0, double2 al1, const cplx_t<R>* diag = nullptr,
                      double2 gamma = {0, 0}) {
    FbOutT<R> o;
    memset(&o, 0, sizeof o);
    o.p = p;
    o.alpha0 = al0;
    o.alpha1 = al1;
    o.diag = diag;
    o.gamma = gamma;
    return o;
}

template <typename R>
static FbJobT<R> job0() {
    FbJobT<R> j;
    memset(&j, 0, sizeof j);
    return j;
}

static const double2 ONE = {1, 0}, MONE = {-1, 0}, I1 = {0, 1}, MI = {0, -1}, ZERO = {0, 0};

template <typename R>
void build_dht_jobs(std::vector<FbJobT<R>>& jobs, const cplx_t<R>* in, int64_t in_m, int64_t in_l, const R* D,
                    int64_t ldd, int64_t d_m, int K, cplx_t<R>* out, int64_t out_m, int64_t out_l, int Ncol,
                    int nmodes, int ncomp) {
    for (int l = 0; l < ncomp; l++)
        for (int m = 0; m < nmodes; m++) {
            FbJobT<R> j = job0<R>();
            j.seg[0] = seg<R>(in + m * in_m + l * in_l, ONE, nullptr, ZERO, D + m * d_m, ldd, 0, 0);
            j.nseg = 1;
            j.out[0] = outp<R>(out + m * out_m + l * out_l, ONE, ZERO);
            j.nout = 1;
            j.K = K;
            j.Ncol = Ncol;
            jobs.push_back(j);
        }
}

// Views of an FB-space array a(nkx, ncol, 0:nmod-1, ncomp): column base of (mode, comp).
template <typename R>
struct FbView {
    const cplx_t<R>* p;
    int64_t nkx, ncol, nmod;
    const cplx_t<R>* at(int m, int c) const { return p + nkx * ncol * (m + nmod * c); }
};
template <typename R>
struct FbViewW {
    cplx_t<R>* p;
    int64_t nkx, ncol, nmod;
    cplx_t<R>* at(int m, int c) const { return p + nkx * ncol * (m + nmod * c); }
};

template <typename R>
void build_rot_jobs(std::vector<FbJobT<R>>& jobs, FbViewW<R> loc, FbView<R> v, const R* Dp, const R* Dm, int nkr,
                    int nkr_loc, int nkO, const R* scale_tab) {
    auto D = [&](const R* B, int m) { return B + (int64_t)nkr * nkr_loc * m; };
    for (int m = 0; m <= nkO; m++) {
        // loc_x = -D+ (i v_y - v_z)[m+1] - D- (i v_y + v_z)[m-1]      fb_math.f90:47-53, :63-69, :76-82
        FbJobT<R> a = job0<R>();
        if (m < nkO) a.seg[a.nseg++] = seg<R>(v.at(m + 1, 1), I1, v.at(m + 1, 2), MONE, D(Dp, m), nkr, 0, 0);
        if (m > 0) a.seg[a.nseg++] = seg<R>(v.at(m - 1, 1), I1, v.at(m - 1, 2), ONE, D(Dm, m), nkr, 0, 0);
        else if (nkO >= 1) a.seg[a.nseg++] = seg<R>(v.at(1, 1), I1, v.at(1, 2), ONE, D(Dm, 0), nkr, 1, 0);
        a.out[0] = outp<R>(loc.at(m, 0), MONE, ZERO);
        a.nout = 1;
        a.K = nkr;
        a.Ncol = nkr_loc;
        // loc_y = -i kx v_z + i W + i V ; loc_z = i kx v_y - W + V     fb_math.f90:42-45, :55-60, :70-75
        FbJobT<R> b = job0<R>();
        if (m < nkO) b.seg[b.nseg++] = seg<R>(v.at(m + 1, 0), ONE, nullptr, ZERO, D(Dp, m), nkr, 0, 0);
        if (m > 0) b.seg[b.nseg++] = seg<R>(v.at(m - 1, 0), ONE, nullptr, ZERO, D(Dm, m), nkr, 0, 1);
        else if (nkO >= 1) b.seg[b.nseg++] = seg<R>(v.at(1, 0), ONE, nullptr, ZERO, D(Dm, 0), nkr, 1, 1);
        b.out[0] = outp<R>(loc.at(m, 1), I1, I1, v.at(m, 2), MI);
        b.out[1] = outp<R>(loc.at(m, 2), MONE, ONE, v.at(m, 1), I1);
        b.nout = 2;
        b.K = nkr;
        b.Ncol = nkr_loc;
        if (scale_tab) {
            const R* sc = scale_tab + (int64_t)v.nkx * nkr_loc * m;
            a.out[0].scale = sc;
            b.out[0].scale = sc;
            b.out[1].scale = sc;
        }
        jobs.push_back(a);
        jobs.push_back(b);
    }
}

template <typename R>
void build_grad_jobs(std::vector<FbJobT<R>>& jobs, FbViewW<R> loc, FbView<R> s, const R* Dp, const R* Dm, int nkr,
                     int nkr_loc, int nkO, int nmod_in) {
    (void)nmod_in;
    auto D = [&](const R* B, int m) { return B + (int64_t)nkr * nkr_loc * m; };
    for (int m = 0; m <= nkO; m++) {
        // fb_math.f90:119-144: loc_x = i kx s[m]; loc_y = -V + W; loc_z = i V + i W
        FbJobT<R> j = job0<R>();
        if (m > 0) j.seg[j.nseg++] = seg<R>(s.at(m - 1, 0), ONE, nullptr, ZERO, D(Dm, m), nkr, 0, 0);
        else if (nkO >= 1) j.seg[j.nseg++] = seg<R>(s.at(1, 0), ONE, nullptr, ZERO, D(Dm, 0), nkr, 1, 0);
        if (m < nkO) j.seg[j.nseg++] = seg<R>(s.at(m + 1, 0), ONE, nullptr, ZERO, D(Dp, m), nkr, 0, 1);
        j.out[0] = outp<R>(loc.at(m, 0), ZERO, ZERO, s.at(m, 0), I1);
        j.out[1] = outp<R>(loc.at(m, 1), MONE, ONE);
        j.out[2] = outp<R>(loc.at(m, 2), I1, I1);
        j.nout = 3;
        j.K = nkr;
        j.Ncol = nkr_loc;
        jobs.push_back(j);
    }
}

// div of v (modes 0..nkO) into out modes 0..nmod_out-1; D+ exists for m < nkO.
template <typename R>
void build_div_jobs(std::vector<FbJobT<R>>& jobs, FbViewW<R> out, FbView<R> v, const R* Dp, const R* Dm, int nkr,
                    int nkr_loc, int nkO, int nmod_out) {
    auto D = [&](const R* B, int m) { return B + (int64_t)nkr * nkr_loc * m; };
    for (int m = 0; m < nmod_out; m++) {
        // fb_math.f90:173-194
        FbJobT<R> j = job0<R>();
        if (m > 0) j.seg[j.nseg++] = seg<R>(v.at(m - 1, 2), I1, v.at(m - 1, 1), MONE, D(Dm, m), nkr, 0, 0);
        else if (nkO >= 1) j.seg[j.nseg++] = seg<R>(v.at(1, 2), I1, v.at(1, 1), MONE, D(Dm, 0), nkr, 1, 0);
        if (m < nkO) j.seg[j.nseg++] = seg<R>(v.at(m + 1, 2), I1, v.at(m + 1, 1), ONE, D(Dp, m), nkr, 0, 0);
        if (m <= nkO) j.out[0] = outp<R>(out.at(m, 0), ONE, ZERO, v.at(m, 0), I1);
        else j.out[0] = outp<R>(out.at(m, 0), ONE, ZERO);
        j.nout = 1;
        j.K = nkr;
        j.Ncol = nkr_loc;
        jobs.push_back(j);
    }
}

// Second half of grad-div: gradient of s (modes 0..nkO+1) into v (modes 0..nkO).
// With poiss != null the outputs take the Poisson-correction epilogue
// v += (graddiv + (gp - gn)*dt_inv)*w2_inv (maxwell_solvers.f90:153-158).
template <typename R>
struct PoissArgs { const cplx_t<R>* gn; const cplx_t<R>* gp; const R* w2; };

template <typename R>
void build_graddiv2_jobs(std::vector<FbJobT<R>>& jobs, FbViewW<R> v, FbView<R> s, const R* Dp, const R* Dm, int nkr,
                         int nkr_loc, int nkO, const PoissArgs<R>* poiss = nullptr) {
    auto D = [&](const R* B, int m) { return B + (int64_t)nkr * nkr_loc * m; };
    for (int m = 0; m <= nkO; m++) {
        // fb_math.f90:266-288
        FbJobT<R> j = job0<R>();
        if (m > 0) j.seg[j.nseg++] = seg<R>(s.at(m - 1, 0), ONE, nullptr, ZERO, D(Dm, m), nkr, 0, 0);
        else if (nkO >= 1) j.seg[j.nseg++] = seg<R>(s.at(1, 0), ONE, nullptr, ZERO, D(Dm, 0), nkr, 1, 0);
        j.seg[j.nseg++] = seg<R>(s.at(m + 1, 0), ONE, nullptr, ZERO, D(Dp, m), nkr, 0, 1);
        j.out[0] = outp<R>(v.at(m, 0), ZERO, ZERO, s.at(m, 0), I1);
        j.out[1] = outp<R>(v.at(m, 1), MONE, ONE);
        j.out[2] = outp<R>(v.at(m, 2), I1, I1);
        j.nout = 3;
        j.K = nkr_loc;
        j.Ncol = nkr_loc;
        if (poiss) {
            const int64_t fc = v.nkx * v.ncol * v.nmod;   // component stride
            const int64_t fm = v.nkx * v.ncol;            // mode stride
            for (int c = 0; c < 3; c++) {
                j.out[c].poiss = 1;
                j.out[c].poiss_n = poiss->gn + m * fm + c * fc;
                j.out[c].poiss_np1 = poiss->gp + m * fm + c * fc;
                j.out[c].poiss_w2 = poiss->w2 + m * fm;
            }
        }
        jobs.push_back(j);
    }
}

// ================================================================== composite device-side operations

template <typename R>
int32_t fb_forward_t(cplx_t<R>* out, cplx_t<R>* in, double leftX, const double* kx_dev, const R* In_dev, int nkx,
                     int nr, int nmodes, int nkr, int ncomp, const R* scale_tab, int fold, const double* leftX_dev) {
    typedef cplx_t<R> C;
    const size_t tmp_elems = (size_t)nkx * nkr * nmodes * ncomp;
    C* tmp = static_cast<C*>(scratch(0, tmp_elems * sizeof(C)));
    C* ph = static_cast<C*>(scratch(1, (size_t)nkx * sizeof(C)));
    if (!tmp || !ph) return ST_ENOMEM;
    if (fold) CH_TRY(ghost_fold<R>(in, nkx, nr + 1, (int64_t)nmodes * ncomp));
    std::vector<FbJobT<R>> jobs;
    // vec(nkx, 0:nr, nmodes, ncomp): row 1 (first real node) is the DHT's r = 1
    build_dht_jobs(jobs, in + nkx, (int64_t)nkx * (nr + 1), (int64_t)nkx * (nr + 1) * nmodes, In_dev, nr,
                   (int64_t)nr * nkr, nr, tmp, (int64_t)nkx * nkr, (int64_t)nkx * nkr * nmodes, nkr, nmodes, ncomp);
    FbGemmArgs g;
    memset(&g, 0, sizeof g);
    g.nx = nkx;
    g.lda = nkx;
    g.ldo = nkx;
    g.kx = kx_dev;
    CH_TRY(launch_fb_gemm<R>(jobs, g, KF_DHT));
    FftArgsT<R> a;
    memset(&a, 0, sizeof a);
    a.ph_at = 2;
    a.ph_s = -1.0;
    a.ph_leftX = leftX;
    a.ph_kx = kx_dev;
    a.ph_leftX_dev = leftX_dev;
    a.ph_table = ph;
    a.in = tmp;
    a.out = out;
    a.nlines = (int64_t)nkr * nmodes * ncomp;
    a.n_inner = nkr * nmodes;
    a.in_r = nkx;
    a.in_q = (int64_t)nkx * nkr * nmodes;
    a.out_r = nkx;
    a.out_q = (int64_t)nkx * nkr * nmodes;
    a.scale = scale_tab;
    a.sign = -1;
    a.N = nkx;
    return launch_fft<R>(a);
}

template <typename R>
int32_t fb_inverse_t(cplx_t<R>* eb, const cplx_t<R>* e_fb, const cplx_t<R>* b_fb, int64_t b_comp_stride, double leftX,
                     const double* kx_dev, const R* Out_dev, int nkx, int nr, int nmodes, int nkr, int ghost,
                     const double* leftX_dev) {
    typedef cplx_t<R> C;
    const size_t tmp_elems = (size_t)nkx * nr * nmodes * 6;
    C* tmp = static_cast<C*>(scratch(0, tmp_elems * sizeof(C)));
    C* ph = static_cast<C*>(scratch(1, (size_t)nkx * sizeof(C)));
    if (!tmp || !ph) return ST_ENOMEM;
    std::vector<FbJobT<R>> jobs;
    const int64_t fbc = (int64_t)nkx * nkr * nmodes;  // e_fb component stride
    build_dht_jobs(jobs, e_fb, (int64_t)nkx * nkr, fbc, Out_dev, nkr, (int64_t)nkr * nr, nkr, tmp, (int64_t)nkx * nr,
                   (int64_t)nkx * nr * nmodes, nr, nmodes, 3);
    build_dht_jobs(jobs, b_fb, (int64_t)nkx * nkr, b_comp_stride, Out_dev, nkr, (int64_t)nkr * nr, nkr,
                   tmp + (size_t)nkx * nr * nmodes * 3, (int64_t)nkx * nr, (int64_t)nkx * nr * nmodes, nr, nmodes, 3);
    FbGemmArgs g;
    memset(&g, 0, sizeof g);
    g.nx = nkx;
    g.lda = nkx;
    g.ldo = nkx;
    g.kx = kx_dev;
    CH_TRY(launch_fb_gemm<R>(jobs, g, KF_DHT));
    FftArgsT<R> a;
    memset(&a, 0, sizeof a);
    a.ph_at = 1;
    a.ph_s = +1.0;
    a.ph_leftX = leftX;
    a.ph_kx = kx_dev;
    a.ph_leftX_dev = leftX_dev;
    a.ph_table = ph;
    a.in = tmp;
    a.out = eb + nkx;  // row 1
    a.nlines = (int64_t)nr * nmodes * 6;
    a.n_inner = nr;
    a.in_r = nkx;
    a.in_q = (int64_t)nkx * nr;
    a.out_r = nkx;
    a.out_q = (int64_t)nkx * (nr + 1);
    a.sign = +1;
    a.ghost = ghost;
    a.nmodes = nmodes;
    a.N = nkx;
    return launch_fft<R>(a);
}

template <typename R>
int32_t fb_filter_t(cplx_t<R>* vec, double leftX, const double* kx_dev, const R* filtr_dev, int nxfilt, int nkx,
                    int nkr, int nmodes, int ncomp) {
    typedef cplx_t<R> C;
    C* ph = static_cast<C*>(scratch(1, (size_t)nkx * sizeof(C)));
    C* phi = static_cast<C*>(scratch(2, (size_t)nkx * sizeof(C)));
    if (!ph || !phi) return ST_ENOMEM;
    CH_TRY(make_phase<R>(ph, kx_dev, leftX, +1.0, nkx));
    hipLaunchKernelGGL((k_phase_inv<R>), blocks_for(nkx, 256), 256, 0, stream(), phi, kx_dev, leftX, nkx);
    CH_LAUNCH_CHECK("k_phase_inv");
    FftArgsT<R> a;
    memset(&a, 0, sizeof a);
    a.in = vec;
    a.out = vec;
    a.nlines = (int64_t)nkr * nmodes * ncomp;
    a.n_inner = 1;
    a.in_r = 0;
    a.in_q = nkx;
    a.out_r = 0;
    a.out_q = nkx;
    a.pre = ph;
    a.post = phi;
    a.sign = +1;
    a.filtr = filtr_dev;
    a.nxfilt = nxfilt < nkx ? nxfilt : nkx;
    a.N = nkx;
    return launch_fft<R>(a);
}

template <typename R>
int32_t fb_op_t(int op, cplx_t<R>* out, const cplx_t<R>* in, const R* Dp, const R* Dm, const double* kx_dev, int nkx,
                int nkr, int nkO, int nkr_loc, const R* scale_tab) {
    typedef cplx_t<R> C;
    std::vector<FbJobT<R>> jobs;
    FbGemmArgs g;
    memset(&g, 0, sizeof g);
    g.nx = nkx;
    g.lda = nkx;
    g.ldo = nkx;
    g.kx = kx_dev;
    const int nmod = nkO + 1;
    switch (op) {
        case FB_ROT:
            build_rot_jobs(jobs, FbViewW<R>{out, nkx, nkr_loc, nmod}, FbView<R>{in, nkx, nkr, nmod}, Dp, Dm, nkr, nkr_loc,
                           nkO, scale_tab);
            return launch_fb_gemm<R>(jobs, g, KF_FBOPS);
        case FB_GRAD:
            build_grad_jobs(jobs, FbViewW<R>{out, nkx, nkr_loc, nmod}, FbView<R>{in, nkx, nkr, nmod}, Dp, Dm, nkr, nkr_loc,
                            nkO, nmod);
            return launch_fb_gemm<R>(jobs, g, KF_FBOPS);
        case FB_DIV:
            build_div_jobs(jobs, FbViewW<R>{out, nkx, nkr_loc, nmod}, FbView<R>{in, nkx, nkr, nmod}, Dp, Dm, nkr, nkr_loc,
                           nkO, nmod);
            return launch_fb_gemm<R>(jobs, g, KF_FBOPS);
        case FB_GRADDIV: {
            if (nkr_loc != nkr) return fail(ST_EUNSUP, "fb_graddiv needs nkr_loc == nkr (got %d, %d)", nkr_loc, nkr);
            C* s = static_cast<C*>(scratch(3, (size_t)nkx * nkr_loc * (nkO + 2) * sizeof(C)));
            if (!s) return ST_ENOMEM;
            build_div_jobs(jobs, FbViewW<R>{s, nkx, nkr_loc, nkO + 2}, FbView<R>{in, nkx, nkr, nmod}, Dp, Dm, nkr, nkr_loc,
                           nkO, nkO + 2);
            CH_TRY(launch_fb_gemm<R>(jobs, g, KF_FBOPS));
            jobs.clear();
            build_graddiv2_jobs(jobs, FbViewW<R>{out, nkx, nkr, nmod}, FbView<R>{s, nkx, nkr_loc, nkO + 2}, Dp, Dm, nkr,
                                nkr_loc, nkO);
            return launch_fb_gemm<R>(jobs, g, KF_FBOPS);
        }
        default:
            return fail(ST_EINVAL, "unknown FB operator %d", op);
    }
}

// j += (graddiv j + (gp - gn)*dt_inv)*w2_inv, one Poisson-correction iteration
// (solvers.py:317-326 with maxwell_solvers.f90:131-164), graddiv fused.
template <typename R>
int32_t fb_graddiv_poisson_t(cplx_t<R>* j, const cplx_t<R>* gn, const cplx_t<R>* gp, double dt_inv, const R* w2,
                             const R* Dp, const R* Dm, const double* kx_dev, int nkx, int nkr, int nkO) {
    typedef cplx_t<R> C;
    std::vector<FbJobT<R>> jobs;
    FbGemmArgs g;
    memset(&g, 0, sizeof g);
    g.nx = nkx;
    g.lda = nkx;
    g.ldo = nkx;
    g.kx = kx_dev;
    g.dt_inv = dt_inv;
    const int nmod = nkO + 1;
    C* s = static_cast<C*>(scratch(3, (size_t)nkx * nkr * (nkO + 2) * sizeof(C)));
    if (!s) return ST_ENOMEM;
    build_div_jobs(jobs, FbViewW<R>{s, nkx, nkr, nkO + 2}, FbView<R>{j, nkx, nkr, nmod}, Dp, Dm, nkr, nkr, nkO, nkO + 2);
    CH_TRY(launch_fb_gemm<R>(jobs, g, KF_FBOPS));
    jobs.clear();
    PoissArgs<R> pa{gn, gp, w2};
    build_graddiv2_jobs(jobs, FbViewW<R>{j, nkx, nkr, nmod}, FbView<R>{s, nkx, nkr, nkO + 2}, Dp, Dm, nkr, nkr, nkO, &pa);
    return launch_fb_gemm<R>(jobs, g, KF_FBOPS);
}


// ---------------------------------------------------------------- envelope (KxShift) operators
//
// fb_math_env.f90: fields carry modes -nkO..nkO (mode m at index m + nkO), the
// coupling matrices D+-(nkr, nkr_loc, -nkO-1:nkO+1) (mode m at m + nkO + 1);
// mode m couples to m-1 and m+1 only where those exist on the grid (no
// reflected mode), and the curl drops the x term of the m-1 coupling as the
// reference does (fb_math_env.f90:141-146: the sum is reset before use).

// rot: loc_x = -D+ (i v_y - v_z)[m+1]; loc_y = -i kx v_z + i S+ + i S-;
// loc_z = i kx v_y - S+ + S-, S+ = D+ v_x[m+1], S- = D- v_x[m-1]
template <typename R>
void build_rot_env_jobs(std::vector<FbJobT<R>>& jobs, FbViewW<R> loc, FbView<R> v, const R* Dp, const R* Dm, int nkr,
                        int nkr_loc, int nkO, const R* scale_tab) {
    auto D = [&](const R* B, int m) { return B + (int64_t)nkr * nkr_loc * (m + nkO + 1); };
    auto X = [&](int m) { return m + nkO; };
    for (int m = -nkO; m <= nkO; m++) {
        FbJobT<R> a = job0<R>();
        if (m < nkO) a.seg[a.nseg++] = seg<R>(v.at(X(m + 1), 1), I1, v.at(X(m + 1), 2), MONE, D(Dp, m), nkr, 0, 0);
        a.out[0] = outp<R>(loc.at(X(m), 0), MONE, ZERO);
        a.nout = 1;
        a.K = nkr;
        a.Ncol = nkr_loc;
        FbJobT<R> b = job0<R>();
        if (m < nkO) b.seg[b.nseg++] = seg<R>(v.at(X(m + 1), 0), ONE, nullptr, ZERO, D(Dp, m), nkr, 0, 0);
        if (m > -nkO) b.seg[b.nseg++] = seg<R>(v.at(X(m - 1), 0), ONE, nullptr, ZERO, D(Dm, m), nkr, 0, 1);
        b.out[0] = outp<R>(loc.at(X(m), 1), I1, I1, v.at(X(m), 2), MI);
        b.out[1] = outp<R>(loc.at(X(m), 2), MONE, ONE, v.at(X(m), 1), I1);
        b.nout = 2;
        b.K = nkr;
        b.Ncol = nkr_loc;
        if (scale_tab) {
            const R* sc = scale_tab + (int64_t)v.nkx * nkr_loc * X(m);
            a.out[0].scale = sc;
            b.out[0].scale = sc;
            b.out[1].scale = sc;
        }
        jobs.push_back(a);
        jobs.push_back(b);
    }
}

// grad: loc_x = i kx s[m]; loc_y = -V + W; loc_z = i V + i W, V = D- s[m-1], W = D+ s[m+1]
// (modes mlo..mhi of s, output modes -nkO..nkO; s of the second half of
// grad-div carries modes -nkO-1..nkO+1 and then every coupling exists)
template <typename R>
void build_grad_env_jobs(std::vector<FbJobT<R>>& jobs, FbViewW<R> loc, FbView<R> s, int s_lo, int s_hi, const R* Dp,
                         const R* Dm, int nkr, int nkr_loc, int K, int nkO, const PoissArgs<R>* poiss = nullptr) {
    auto D = [&](const R* B, int m) { return B + (int64_t)nkr * nkr_loc * (m + nkO + 1); };
    for (int m = -nkO; m <= nkO; m++) {
        FbJobT<R> j = job0<R>();
        if (m - 1 >= s_lo) j.seg[j.nseg++] = seg<R>(s.at(m - 1 - s_lo, 0), ONE, nullptr, ZERO, D(Dm, m), nkr, 0, 0);
        if (m + 1 <= s_hi) j.seg[j.nseg++] = seg<R>(s.at(m + 1 - s_lo, 0), ONE, nullptr, ZERO, D(Dp, m), nkr, 0, 1);
        j.out[0] = outp<R>(loc.at(m + nkO, 0), ZERO, ZERO, s.at(m - s_lo, 0), I1);
        j.out[1] = outp<R>(loc.at(m + nkO, 1), MONE, ONE);
        j.out[2] = outp<R>(loc.at(m + nkO, 2), I1, I1);
        j.nout = 3;
        j.K = K;
        j.Ncol = nkr_loc;
        if (poiss) {
            const int64_t fc = loc.nkx * loc.ncol * loc.nmod, fm = loc.nkx * loc.ncol;
            for (int c = 0; c < 3; c++) {
                j.out[c].poiss = 1;
                j.out[c].poiss_n = poiss->gn + (m + nkO) * fm + c * fc;
                j.out[c].poiss_np1 = poiss->gp + (m + nkO) * fm + c * fc;
                j.out[c].poiss_w2 = poiss->w2 + (m + nkO) * fm;
            }
        }
        jobs.push_back(j);
    }
}

// div into output modes olo..ohi: out = i kx v_x[m] (m on the grid)
// - D- (v_y - i v_z)[m-1] + D+ (v_y + i v_z)[m+1]
template <typename R>
void build_div_env_jobs(std::vector<FbJobT<R>>& jobs, FbViewW<R> out, int olo, int ohi, FbView<R> v, const R* Dp,
                        const R* Dm, int nkr, int nkr_loc, int nkO) {
    auto D = [&](const R* B, int m) { return B + (int64_t)nkr * nkr_loc * (m + nkO + 1); };
    for (int m = olo; m <= ohi; m++) {
        FbJobT<R> j = job0<R>();
        if (m > -nkO) j.seg[j.nseg++] = seg<R>(v.at(m - 1 + nkO, 2), I1, v.at(m - 1 + nkO, 1), MONE, D(Dm, m), nkr, 0, 0);
        if (m < nkO) j.seg[j.nseg++] = seg<R>(v.at(m + 1 + nkO, 2), I1, v.at(m + 1 + nkO, 1), ONE, D(Dp, m), nkr, 0, 0);
        if (m >= -nkO && m <= nkO) j.out[0] = outp<R>(out.at(m - olo, 0), ONE, ZERO, v.at(m + nkO, 0), I1);
        else j.out[0] = outp<R>(out.at(m - olo, 0), ONE, ZERO);
        j.nout = 1;
        j.K = nkr;
        j.Ncol = nkr_loc;
        jobs.push_back(j);
    }
}

template <typename R>
int32_t fb_op_env_t(int op, cplx_t<R>* out, const cplx_t<R>* in, const R* Dp, const R* Dm, const double* kx_dev,
                    int nkx, int nkr, int nkO, int nkr_loc, const R* scale_tab, const PoissArgs<R>* poiss,
                    double dt_inv) {
    typedef cplx_t<R> C;
    std::vector<FbJobT<R>> jobs;
    FbGemmArgs g;
    memset(&g, 0, sizeof g);
    g.nx = nkx;
    g.lda = nkx;
    g.ldo = nkx;
    g.kx = kx_dev;
    g.dt_inv = dt_inv;
    const int nm = 2 * nkO + 1;
    switch (op) {
        case FB_ROT:
            build_rot_env_jobs(jobs, FbViewW<R>{out, nkx, nkr_loc, nm}, FbView<R>{in, nkx, nkr, nm}, Dp, Dm, nkr,
                               nkr_loc, nkO, scale_tab);
            return launch_fb_gemm<R>(jobs, g, KF_FBOPS);
        case FB_GRAD:
            build_grad_env_jobs(jobs, FbViewW<R>{out, nkx, nkr_loc, nm}, FbView<R>{in, nkx, nkr, nm}, -nkO, nkO, Dp, Dm,
                                nkr, nkr_loc, nkr, nkO);
            return launch_fb_gemm<R>(jobs, g, KF_FBOPS);
        case FB_DIV:
            build_div_env_jobs(jobs, FbViewW<R>{out, nkx, nkr_loc, nm}, -nkO, nkO, FbView<R>{in, nkx, nkr, nm}, Dp, Dm,
                               nkr, nkr_loc, nkO);
            return launch_fb_gemm<R>(jobs, g, KF_FBOPS);
        case FB_GRADDIV: {
            if (nkr_loc != nkr) return fail(ST_EUNSUP, "fb_graddiv_env needs nkr_loc == nkr (got %d, %d)", nkr_loc, nkr);
            C* s = static_cast<C*>(scratch(3, (size_t)nkx * nkr * (nm + 2) * sizeof(C)));
            if (!s) return ST_ENOMEM;
            build_div_env_jobs(jobs, FbViewW<R>{s, nkx, nkr, nm + 2}, -nkO - 1, nkO + 1, FbView<R>{in, nkx, nkr, nm},
                               Dp, Dm, nkr, nkr, nkO);
            CH_TRY(launch_fb_gemm<R>(jobs, g, KF_FBOPS));
            jobs.clear();
            build_grad_env_jobs(jobs, FbViewW<R>{out, nkx, nkr, nm}, FbView<R>{s, nkx, nkr, nm + 2}, -nkO - 1, nkO + 1,
                                Dp, Dm, nkr, nkr, nkr, nkO, poiss);
            return launch_fb_gemm<R>(jobs, g, KF_FBOPS);
        }
        default:
            return fail(ST_EINVAL, "unknown FB operator %d", op);
    }
}

#define CH_FB_INSTANTIATE(R)                                                                                          \
    template int32_t fb_forward_t<R>(cplx_t<R>*, cplx_t<R>*, double, const double*, const R*, int, int, int, int, int, \
                                     const R*, int, const double*);                                                   \
    template int32_t fb_inverse_t<R>(cplx_t<R>*, const cplx_t<R>*, const cplx_t<R>*, int64_t, double, const double*,  \
                                     const R*, int, int, int, int, int, const double*);                               \
    template int32_t fb_filter_t<R>(cplx_t<R>*, double, const double*, const R*, int, int, int, int, int);            \
    template int32_t fb_op_t<R>(int, cplx_t<R>*, const cplx_t<R>*, const R*, const R*, const double*, int, int, int,  \
                                int, const R*);                                                                       \
    template int32_t fb_graddiv_poisson_t<R>(cplx_t<R>*, const cplx_t<R>*, const cplx_t<R>*, double, const R*,         \
                                             const R*, const R*, const double*, int, int, int);                       \
    template int32_t fb_op_env_t<R>(int, cplx_t<R>*, const cplx_t<R>*, const R*, const R*, const double*, int, int,   \
                                    int, int, const R*, const PoissArgs<R>*, double);
CH_FB_INSTANTIATE(double)
CH_FB_INSTANTIATE(float)
#undef CH_FB_INSTANTIATE

}  // namespace chimera

using namespace chimera;

extern "C" {

int32_t chimera_fft_length_check(int32_t n) { return fft_length_check(n); }

int32_t chimera_set_mixed_gemm(int32_t form) {
    if (form != MIXED_GEMM_BF16X6 && form != MIXED_GEMM_FP32)
        return fail(ST_EINVAL, "chimera_set_mixed_gemm: form %d (0 = split-BF16, 1 = FP32 MFMA)", form);
    g_mixed_gemm.store(form);
    return ST_OK;
}

int32_t chimera_get_mixed_gemm(void) { return g_mixed_gemm.load(); }

int32_t chimera_set_autotune(int32_t on, const char* path) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_autotune.store(on ? 1 : 0);
    const std::string p = path ? path : "";
    if (p != g_tune_path) {
        g_tune_path = p;
        g_tune_loaded = false;
    }
    return ST_OK;
}

int32_t chimera_autotune_entries(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return (int32_t)g_tuned.size();
}

int32_t chimera_fb_vec_in(double* vec_fb, const double* vec, double leftX, const double* kx, const double* In,
                          int32_t nkx, int32_t nr, int32_t nkO, int32_t nkr) {
    if (nkx < 1 || nr < 1 || nkO < 1 || nkr < 1) return fail(ST_EINVAL, "fb_vec_in: bad sizes");
    CH_TRY(fft_length_check(nkx));
    Stage st;
    double2* o = reinterpret_cast<double2*>(st.out(vec_fb, 2 * (size_t)nkx * nkr * nkO * 3));
    const double2* v = reinterpret_cast<const double2*>(st.in(vec, 2 * (size_t)nkx * (nr + 1) * nkO * 3));
    const double* k = st.in(kx, nkx);
    const double* M = st.in(In, (size_t)nr * nkr * nkO);
    CH_TRY(st.status());
    CH_TRY(fb_forward(o, const_cast<double2*>(v), leftX, k, M, nkx, nr, nkO, nkr, 3, nullptr, 0));  // fold 0: read only
    return st.finish();
}

int32_t chimera_fb_scl_in(double* scl_fb, const double* scl, double leftX, const double* kx, const double* In,
                          int32_t nkx, int32_t nr, int32_t nkO, int32_t nkr) {
    if (nkx < 1 || nr < 1 || nkO < 1 || nkr < 1) return fail(ST_EINVAL, "fb_scl_in: bad sizes");
    CH_TRY(fft_length_check(nkx));
    Stage st;
    double2* o = reinterpret_cast<double2*>(st.out(scl_fb, 2 * (size_t)nkx * nkr * nkO));
    const double2* v = reinterpret_cast<const double2*>(st.in(scl, 2 * (size_t)nkx * (nr + 1) * nkO));
    const double* k = st.in(kx, nkx);
    const double* M = st.in(In, (size_t)nr * nkr * nkO);
    CH_TRY(st.status());
    CH_TRY(fb_forward(o, const_cast<double2*>(v), leftX, k, M, nkx, nr, nkO, nkr, 1, nullptr, 0));  // fold 0: read only
    return st.finish();
}

int32_t chimera_fb_eb_out(double* eb_spc, const double* e_fb, const double* b_fb, double leftX, const double* kx,
                          const double* Out, int32_t nkx, int32_t nr, int32_t nkO, int32_t nkr) {
    if (nkx < 1 || nr < 1 || nkO < 1 || nkr < 1) return fail(ST_EINVAL, "fb_eb_out: bad sizes");
    CH_TRY(fft_length_check(nkx));
    Stage st;
    double2* eb = reinterpret_cast<double2*>(st.out(eb_spc, 2 * (size_t)nkx * (nr + 1) * nkO * 6));
    const double2* e = reinterpret_cast<const double2*>(st.in(e_fb, 2 * (size_t)nkx * nkr * nkO * 6));
    const double2* b = reinterpret_cast<const double2*>(st.in(b_fb, 2 * (size_t)nkx * nkr * nkO * 3));
    const double* k = st.in(kx, nkx);
    const double* M = st.in(Out, (size_t)nkr * nr * nkO);
    CH_TRY(st.status());
    CH_TRY(fb_inverse(eb, e, b, (int64_t)nkx * nkr * nkO, leftX, k, M, nkx, nr, nkO, nkr, 1));
    return st.finish();
}

int32_t chimera_fb_filtr(double* vec, double leftX, const double* kx, const double* filtr, int32_t modefilt,
                         int32_t nkx, int32_t nkr, int32_t nkO, int32_t nxfilt) {
    if (modefilt != 0)
        return fail(ST_EUNSUP, "fb_filtr: modefilt=%d; only 0 ('left') is defined (the reference's other modes index "
                               "the profile out of its bounds)", modefilt);
    if (nkx < 1 || nkr < 1 || nkO < 1 || nxfilt < 0) return fail(ST_EINVAL, "fb_filtr: bad sizes");
    CH_TRY(fft_length_check(nkx));
    Stage st;
    double2* v = reinterpret_cast<double2*>(st.inout(vec, 2 * (size_t)nkx * nkr * nkO * 3));
    const double* k = st.in(kx, nkx);
    const double* f = st.in(filtr, (size_t)(nxfilt > 0 ? nxfilt : 0));
    CH_TRY(st.status());
    CH_TRY(fb_filter(v, leftX, k, nxfilt > 0 ? f : nullptr, nxfilt, nkx, nkr, nkO, 3));
    return st.finish();
}

static int32_t fb_op_entry(int op, double* out, const double* in, const double* DpS2S, const double* DmS2S,
                           const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc, int in_comp,
                           int out_comp) {
    if (nkx < 1 || nkr < 1 || nkO < 0 || nkr_loc < 1) return fail(ST_EINVAL, "FB operator: bad sizes");
    if (nkr_loc > nkr) return fail(ST_EINVAL, "FB operator: nkr_loc=%d exceeds nkr=%d", nkr_loc, nkr);
    Stage st;
    const size_t nin = (size_t)nkx * nkr * (nkO + 1) * in_comp;
    const size_t nout = (size_t)nkx * nkr_loc * (nkO + 1) * out_comp;
    double2* o;
    const double2* i;
    if (op == FB_GRADDIV) {
        o = reinterpret_cast<double2*>(st.inout(out, 2 * nin));
        i = o;
    } else {
        o = reinterpret_cast<double2*>(st.out(out, 2 * nout));
        i = reinterpret_cast<const double2*>(st.in(in, 2 * nin));
    }
    const size_t nd = (size_t)nkr * nkr_loc * (nkO + 2);
    const double* Dp = st.in(DpS2S, nd);
    const double* Dm = st.in(DmS2S, nd);
    const double* k = st.in(kx, nkx);
    CH_TRY(st.status());
    CH_TRY(fb_op(op, o, i, Dp, Dm, k, nkx, nkr, nkO, nkr_loc, nullptr));
    return st.finish();
}

int32_t chimera_fb_rot(double* vec_fb_loc, const double* vec_fb, const double* DpS2S, const double* DmS2S,
                       const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc) {
    return fb_op_entry(FB_ROT, vec_fb_loc, vec_fb, DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr_loc, 3, 3);
}

int32_t chimera_fb_grad(double* vec_fb_loc, const double* scl_fb, const double* DpS2S, const double* DmS2S,
                        const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc) {
    return fb_op_entry(FB_GRAD, vec_fb_loc, scl_fb, DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr_loc, 1, 3);
}

int32_t chimera_fb_div(double* scl_fb_loc, const double* vec_fb, const double* DpS2S, const double* DmS2S,
                       const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc) {
    return fb_op_entry(FB_DIV, scl_fb_loc, vec_fb, DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr_loc, 3, 1);
}

int32_t chimera_fb_graddiv(double* vec_fb, const double* DpS2S, const double* DmS2S, const double* kx, int32_t nkx,
                           int32_t nkr, int32_t nkO, int32_t nkr_loc) {
    return fb_op_entry(FB_GRADDIV, vec_fb, vec_fb, DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr_loc, 3, 3);
}

static int32_t fb_op_env_entry(int op, double* out, const double* in, const double* DpS2S, const double* DmS2S,
                               const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc, int in_comp,
                               int out_comp) {
    if (nkx < 1 || nkr < 1 || nkO < 0 || nkr_loc < 1) return fail(ST_EINVAL, "FB envelope operator: bad sizes");
    if (nkr_loc > nkr) return fail(ST_EINVAL, "FB envelope operator: nkr_loc=%d exceeds nkr=%d", nkr_loc, nkr);
    const int nm = 2 * nkO + 1;
    Stage st;
    const size_t nin = (size_t)nkx * nkr * nm * in_comp;
    const size_t nout = (size_t)nkx * nkr_loc * nm * out_comp;
    double2* o;
    const double2* i;
    if (op == FB_GRADDIV) {
        o = reinterpret_cast<double2*>(st.inout(out, 2 * nin));
        i = o;
    } else {
        o = reinterpret_cast<double2*>(st.out(out, 2 * nout));
        i = reinterpret_cast<const double2*>(st.in(in, 2 * nin));
    }
    const size_t nd = (size_t)nkr * nkr_loc * (nm + 2);
    const double* Dp = st.in(DpS2S, nd);
    const double* Dm = st.in(DmS2S, nd);
    const double* k = st.in(kx, nkx);
    CH_TRY(st.status());
    CH_TRY(fb_op_env_t<double>(op, o, i, Dp, Dm, k, nkx, nkr, nkO, nkr_loc, nullptr, nullptr, 0.0));
    return st.finish();
}

int32_t chimera_fb_rot_env(double* vec_fb_loc, const double* vec_fb, const double* DpS2S, const double* DmS2S,
                           const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc) {
    return fb_op_env_entry(FB_ROT, vec_fb_loc, vec_fb, DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr_loc, 3, 3);
}

int32_t chimera_fb_grad_env(double* vec_fb_loc, const double* scl_fb, const double* DpS2S, const double* DmS2S,
                            const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc) {
    return fb_op_env_entry(FB_GRAD, vec_fb_loc, scl_fb, DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr_loc, 1, 3);
}

int32_t chimera_fb_div_env(double* scl_fb_loc, const double* vec_fb, const double* DpS2S, const double* DmS2S,
                           const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc) {
    return fb_op_env_entry(FB_DIV, scl_fb_loc, vec_fb, DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr_loc, 3, 1);
}

int32_t chimera_fb_graddiv_env(double* vec_fb, const double* DpS2S, const double* DmS2S, const double* kx,
                               int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc) {
    return fb_op_env_entry(FB_GRADDIV, vec_fb, vec_fb, DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr_loc, 3, 3);
}

/* ---------------------------------------------------------------- resident tier (device pointers) */

int32_t chimera_fb_forward(double* out, double* in, double leftX, const double* kx, const double* In,
                           const double* scale, int32_t nkx, int32_t nr, int32_t nmodes, int32_t nkr, int32_t ncomp,
                           int32_t fold, const double* frame) {
    if (nkx < 1 || nr < 1 || nmodes < 1 || nkr < 1 || ncomp < 1) return fail(ST_EINVAL, "fb_forward: bad sizes");
    CH_TRY(fft_length_check(nkx));
    if (!is_device_ptr(out) || !is_device_ptr(in) || !is_device_ptr(kx) || !is_device_ptr(In) ||
        (scale && !is_device_ptr(scale)))
        return fail(ST_EINVAL, "fb_forward: resident entry points take device pointers");
    return fb_forward_t<double>(reinterpret_cast<double2*>(out), reinterpret_cast<double2*>(in), leftX, kx, In, nkx,
                                nr, nmodes, nkr, ncomp, scale, fold, frame);
}

int32_t chimera_fb_inverse(double* eb_spc, const double* e_fb, const double* b_fb, double leftX, const double* kx,
                           const double* Out, int32_t nkx, int32_t nr, int32_t nmodes, int32_t nkr, int32_t ghost,
                           const double* frame) {
    if (nkx < 1 || nr < 1 || nmodes < 1 || nkr < 1 || ghost < 1 || ghost > 4)
        return fail(ST_EINVAL, "fb_inverse: bad sizes or ghost mode");
    CH_TRY(fft_length_check(nkx));
    if (!is_device_ptr(eb_spc) || !is_device_ptr(e_fb) || !is_device_ptr(b_fb) || !is_device_ptr(kx) ||
        !is_device_ptr(Out))
        return fail(ST_EINVAL, "fb_inverse: resident entry points take device pointers");
    return fb_inverse_t<double>(reinterpret_cast<double2*>(eb_spc), reinterpret_cast<const double2*>(e_fb),
                                reinterpret_cast<const double2*>(b_fb), (int64_t)nkx * nkr * nmodes, leftX, kx, Out,
                                nkx, nr, nmodes, nkr, ghost, frame);
}

int32_t chimera_fb_rot_scaled(double* vec_fb_loc, const double* vec_fb, const double* DpS2S, const double* DmS2S,
                              const double* kx, const double* scale, int32_t nkx, int32_t nkr, int32_t nkO) {
    if (nkx < 1 || nkr < 1 || nkO < 0) return fail(ST_EINVAL, "fb_rot_scaled: bad sizes");
    if (!is_device_ptr(vec_fb_loc) || !is_device_ptr(vec_fb) || !is_device_ptr(scale))
        return fail(ST_EINVAL, "fb_rot_scaled: resident entry points take device pointers");
    return fb_op(FB_ROT, reinterpret_cast<double2*>(vec_fb_loc), reinterpret_cast<const double2*>(vec_fb), DpS2S,
                 DmS2S, kx, nkx, nkr, nkO, nkr, scale);
}

int32_t chimera_fb_graddiv_poisson(double* j_fb, const double* grad_rho_n_fb, const double* grad_rho_np1_fb,
                                   double dt_inv, const double* w2_inv, const double* DpS2S, const double* DmS2S,
                                   const double* kx, int32_t nkx, int32_t nkr, int32_t nkO) {
    if (nkx < 1 || nkr < 1 || nkO < 0) return fail(ST_EINVAL, "fb_graddiv_poisson: bad sizes");
    if (!is_device_ptr(j_fb) || !is_device_ptr(grad_rho_n_fb) || !is_device_ptr(grad_rho_np1_fb) ||
        !is_device_ptr(w2_inv))
        return fail(ST_EINVAL, "fb_graddiv_poisson: resident entry points take device pointers");
    return fb_graddiv_poisson_t<double>(reinterpret_cast<double2*>(j_fb), reinterpret_cast<const double2*>(grad_rho_n_fb),
                              reinterpret_cast<const double2*>(grad_rho_np1_fb), dt_inv, w2_inv, DpS2S, DmS2S, kx,
                              nkx, nkr, nkO);
}

int32_t chimera_fb_op(int32_t op, double* out, const double* in, const double* DpS2S, const double* DmS2S,
                      const double* kx, int32_t nkx, int32_t nkr, int32_t nkO) {
    if (nkx < 1 || nkr < 1 || nkO < 0 || op < 0 || op > 3) return fail(ST_EINVAL, "fb_op: bad arguments");
    if (!is_device_ptr(out) || !is_device_ptr(in)) return fail(ST_EINVAL, "fb_op: resident entry points take device pointers");
    return fb_op(op, reinterpret_cast<double2*>(out), reinterpret_cast<const double2*>(in), DpS2S, DmS2S, kx, nkx, nkr,
                 nkO, nkr, nullptr);
}

int32_t chimera_fb_filter(double* vec, double leftX, const double* kx, const double* filtr, int32_t nxfilt,
                          int32_t nkx, int32_t nkr, int32_t nmodes, int32_t ncomp) {
    if (nkx < 1 || nkr < 1 || nmodes < 1 || ncomp < 1 || nxfilt < 0) return fail(ST_EINVAL, "fb_filter: bad sizes");
    CH_TRY(fft_length_check(nkx));
    if (!is_device_ptr(vec) || !is_device_ptr(kx) || (nxfilt > 0 && !is_device_ptr(filtr)))
        return fail(ST_EINVAL, "fb_filter: resident entry points take device pointers");
    return fb_filter(reinterpret_cast<double2*>(vec), leftX, kx, nxfilt > 0 ? filtr : nullptr, nxfilt, nkx, nkr,
                     nmodes, ncomp);
}

/* ---------------------------------------------------------------- resident tier, complex64 (mixed precision) */

int32_t chimera_fb_forward_c8(float* out, float* in, double leftX, const double* kx, const float* In,
                              const float* scale, int32_t nkx, int32_t nr, int32_t nmodes, int32_t nkr, int32_t ncomp,
                              int32_t fold, const double* frame) {
    if (nkx < 1 || nr < 1 || nmodes < 1 || nkr < 1 || ncomp < 1) return fail(ST_EINVAL, "fb_forward_c8: bad sizes");
    CH_TRY(fft_length_check(nkx));
    if (!is_device_ptr(out) || !is_device_ptr(in) || !is_device_ptr(kx) || !is_device_ptr(In) ||
        (scale && !is_device_ptr(scale)))
        return fail(ST_EINVAL, "fb_forward_c8: resident entry points take device pointers");
    return fb_forward_t<float>(reinterpret_cast<float2*>(out), reinterpret_cast<float2*>(in), leftX, kx, In, nkx, nr,
                               nmodes, nkr, ncomp, scale, fold, frame);
}

int32_t chimera_fb_inverse_c8(float* eb_spc, const float* e_fb, const float* b_fb, double leftX, const double* kx,
                              const float* Out, int32_t nkx, int32_t nr, int32_t nmodes, int32_t nkr, int32_t ghost,
                              const double* frame) {
    if (nkx < 1 || nr < 1 || nmodes < 1 || nkr < 1 || ghost < 1 || ghost > 4)
        return fail(ST_EINVAL, "fb_inverse_c8: bad sizes or ghost mode");
    CH_TRY(fft_length_check(nkx));
    if (!is_device_ptr(eb_spc) || !is_device_ptr(e_fb) || !is_device_ptr(b_fb) || !is_device_ptr(kx) ||
        !is_device_ptr(Out))
        return fail(ST_EINVAL, "fb_inverse_c8: resident entry points take device pointers");
    return fb_inverse_t<float>(reinterpret_cast<float2*>(eb_spc), reinterpret_cast<const float2*>(e_fb),
                               reinterpret_cast<const float2*>(b_fb), (int64_t)nkx * nkr * nmodes, leftX, kx, Out,
                               nkx, nr, nmodes, nkr, ghost, frame);
}

int32_t chimera_fb_rot_scaled_c8(float* vec_fb_loc, const float* vec_fb, const float* DpS2S, const float* DmS2S,
                                 const double* kx, const float* scale, int32_t nkx, int32_t nkr, int32_t nkO) {
    if (nkx < 1 || nkr < 1 || nkO < 0) return fail(ST_EINVAL, "fb_rot_scaled_c8: bad sizes");
    if (!is_device_ptr(vec_fb_loc) || !is_device_ptr(vec_fb) || !is_device_ptr(scale))
        return fail(ST_EINVAL, "fb_rot_scaled_c8: resident entry points take device pointers");
    return fb_op_t<float>(FB_ROT, reinterpret_cast<float2*>(vec_fb_loc), reinterpret_cast<const float2*>(vec_fb),
                          DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr, scale);
}

int32_t chimera_fb_graddiv_poisson_c8(float* j_fb, const float* grad_rho_n_fb, const float* grad_rho_np1_fb,
                                      double dt_inv, const float* w2_inv, const float* DpS2S, const float* DmS2S,
                                      const double* kx, int32_t nkx, int32_t nkr, int32_t nkO) {
    if (nkx < 1 || nkr < 1 || nkO < 0) return fail(ST_EINVAL, "fb_graddiv_poisson_c8: bad sizes");
    if (!is_device_ptr(j_fb) || !is_device_ptr(grad_rho_n_fb) || !is_device_ptr(grad_rho_np1_fb) ||
        !is_device_ptr(w2_inv))
        return fail(ST_EINVAL, "fb_graddiv_poisson_c8: resident entry points take device pointers");
    return fb_graddiv_poisson_t<float>(reinterpret_cast<float2*>(j_fb), reinterpret_cast<const float2*>(grad_rho_n_fb),
                                       reinterpret_cast<const float2*>(grad_rho_np1_fb), dt_inv, w2_inv, DpS2S, DmS2S,
                                       kx, nkx, nkr, nkO);
}

int32_t chimera_fb_op_c8(int32_t op, float* out, const float* in, const float* DpS2S, const float* DmS2S,
                         const double* kx, int32_t nkx, int32_t nkr, int32_t nkO) {
    if (nkx < 1 || nkr < 1 || nkO < 0 || op < 0 || op > 3) return fail(ST_EINVAL, "fb_op_c8: bad arguments");
    if (!is_device_ptr(out) || !is_device_ptr(in))
        return fail(ST_EINVAL, "fb_op_c8: resident entry points take device pointers");
    return fb_op_t<float>(op, reinterpret_cast<float2*>(out), reinterpret_cast<const float2*>(in), DpS2S, DmS2S, kx,
                          nkx, nkr, nkO, nkr, nullptr);
}

int32_t chimera_fb_filter_c8(float* vec, double leftX, const double* kx, const float* filtr, int32_t nxfilt,
                             int32_t nkx, int32_t nkr, int32_t nmodes, int32_t ncomp) {
    if (nkx < 1 || nkr < 1 || nmodes < 1 || ncomp < 1 || nxfilt < 0) return fail(ST_EINVAL, "fb_filter_c8: bad sizes");
    CH_TRY(fft_length_check(nkx));
    if (!is_device_ptr(vec) || !is_device_ptr(kx) || (nxfilt > 0 && !is_device_ptr(filtr)))
        return fail(ST_EINVAL, "fb_filter_c8: resident entry points take device pointers");
    return fb_filter_t<float>(reinterpret_cast<float2*>(vec), leftX, kx, nxfilt > 0 ? filtr : nullptr, nxfilt, nkx,
                              nkr, nmodes, ncomp);
}

}  // extern "C"
