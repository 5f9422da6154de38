// fft.hpp -- batched complex FFT along x (R = double: complex128, R = float:
// complex64), one block of TPB threads (256 by default) per TOT = TPB*EPT
// points of LDS (one line of length N = TOT, or TOT/N lines).
//
// Stockham autosort formulation: every pass reads R points at stride N/R,
// twiddles them, does a radix-R DFT and writes them back in natural order, so
// no bit reversal is needed.  All points of a pass are held in registers
// between one barrier and the next, so a single LDS buffer suffices and a line
// of 8192 points (128 KiB) fits in one workgroup.  Radix 4, with one radix-2
// pass first when log2 N is odd.  Twiddles come from a table tw[k] =
// exp(-2 pi i k/N) built on the host in long double and rounded to R.  Sign conventions are
// FFTW's: FORWARD = exp(-2 pi i jk/N), BACKWARD = exp(+...), both unnormalised.
#pragma once

#include "common.hpp"

namespace chimera {

constexpr int FFT_THREADS = 256;

template <typename C>
__device__ __forceinline__ C tw_get(const C* __restrict__ tw, int idx, int sign) {
    const C w = tw[idx];
    return sign < 0 ? w : cconj(w);
}

template <int EPT, typename R, int TPB = FFT_THREADS>
__device__ __forceinline__ void fft_pass_r4(cplx_t<R>* buf, int N, int Ns, int sign, const cplx_t<R>* __restrict__ tw) {
    typedef cplx_t<R> C;
    constexpr int NB = EPT / 4;
    const int tid = threadIdx.x;
    const int q4 = N >> 2;
    C v[NB][4];
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int b = tid + TPB * q;
        const int line = b / q4, j = b - line * q4;
        const C* src = buf + line * N + j;
#pragma unroll
        for (int r = 0; r < 4; r++) v[q][r] = src[r * q4];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int b = tid + TPB * q;
        const int line = b / q4, j = b - line * q4;
        const int k = j % Ns;
        if (Ns > 1) {
            const int step = N / (Ns * 4);
#pragma unroll
            for (int r = 1; r < 4; r++) v[q][r] = cmul(v[q][r], tw_get(tw, k * r * step, sign));
        }
        const C a0 = cadd(v[q][0], v[q][2]), a1 = csub(v[q][0], v[q][2]);
        const C a2 = cadd(v[q][1], v[q][3]);
        C a3 = csub(v[q][1], v[q][3]);
        a3 = sign < 0 ? cmake<R>(a3.y, -a3.x) : cmake<R>(-a3.y, a3.x);  // times -i (forward) or +i
        C* dst = buf + line * N + (j / Ns) * Ns * 4 + k;
        dst[0] = cadd(a0, a2);
        dst[Ns] = cadd(a1, a3);
        dst[2 * Ns] = csub(a0, a2);
        dst[3 * Ns] = csub(a1, a3);
    }
    __syncthreads();
}

template <int EPT, typename R, int TPB = FFT_THREADS>
__device__ __forceinline__ void fft_pass_r2(cplx_t<R>* buf, int N, int Ns, int sign, const cplx_t<R>* __restrict__ tw) {
    typedef cplx_t<R> C;
    constexpr int NB = EPT / 2;
    const int tid = threadIdx.x;
    const int q2 = N >> 1;
    C v[NB][2];
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int b = tid + TPB * q;
        const int line = b / q2, j = b - line * q2;
        v[q][0] = buf[line * N + j];
        v[q][1] = buf[line * N + j + q2];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int b = tid + TPB * q;
        const int line = b / q2, j = b - line * q2;
        const int k = j % Ns;
        if (Ns > 1) v[q][1] = cmul(v[q][1], tw_get(tw, k * (N / (Ns * 2)), sign));
        C* dst = buf + line * N + (j / Ns) * Ns * 2 + k;
        dst[0] = cadd(v[q][0], v[q][1]);
        dst[Ns] = csub(v[q][0], v[q][1]);
    }
    __syncthreads();
}

// In-place transform of the TOT/N lines held in buf (N a power of two).
template <int EPT, typename R, int TPB = FFT_THREADS>
__device__ __forceinline__ void fft_lds(cplx_t<R>* buf, int N, int log2N, int sign, const cplx_t<R>* __restrict__ tw) {
    int Ns = 1;
    if (log2N & 1) {
        fft_pass_r2<EPT, R, TPB>(buf, N, Ns, sign, tw);
        Ns = 2;
    }
    while (Ns < N) {
        fft_pass_r4<EPT, R, TPB>(buf, N, Ns, sign, tw);
        Ns *= 4;
    }
}

// Twiddle table exp(-2 pi i k / N), k < N, cached per N on the device.
const double2* fft_twiddles(int N);
const float2* fft_twiddles_f(int N);

}  // namespace chimera
