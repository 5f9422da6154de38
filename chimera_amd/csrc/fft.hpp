// fft.hpp -- batched complex FFT along x (R = double: complex128, R = float:
// complex64), one block of TPB threads (256 by default) per TOT = TPB*EPT
// points of LDS (one line of length N = TOT, or TOT/N lines).
//
// Stockham autosort formulation: every pass reads R points at stride N/R,
// twiddles them, does a radix-R DFT and writes them back in natural order, so
// no bit reversal is needed.  All points of a pass are held in registers
// between one barrier and the next, so a single LDS buffer suffices and a line
// of 8192 points (128 KiB) fits in one workgroup.  Radix 8 (with one radix-4 or
// radix-2 pass last) when a thread holds a multiple of eight points, else radix
// 4 with one radix-2 pass first when log2 N is odd.  Twiddles come from a table tw[k] =
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

// Radix-8 pass (EPT a multiple of 8): eight points at stride N/8, twiddled, an
// 8-point DFT as two 4-point DFTs of the even and odd points combined with the
// eighth roots of unity, written back in natural order.
template <int EPT, typename R, int TPB = FFT_THREADS>
__device__ __forceinline__ void fft_pass_r8(cplx_t<R>* buf, int N, int Ns, int sign, const cplx_t<R>* __restrict__ tw) {
    typedef cplx_t<R> C;
    constexpr int NB = EPT / 8;
    const int tid = threadIdx.x;
    const int q8 = N >> 3;
    const R h = (R)0.70710678118654752440;    // 1/sqrt(2)
    C v[NB][8];
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int b = tid + TPB * q;
        const int line = b / q8, j = b - line * q8;
        const C* src = buf + line * N + j;
#pragma unroll
        for (int r = 0; r < 8; r++) v[q][r] = src[r * q8];
    }
    __syncthreads();
    // s*i*z: the 4-point DFT's and W8's imaginary unit with the transform's sign
    auto si = [sign](C z) __attribute__((always_inline)) {
        return sign < 0 ? cmake<R>(z.y, -z.x) : cmake<R>(-z.y, z.x);
    };
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int b = tid + TPB * q;
        const int line = b / q8, j = b - line * q8;
        const int k = j % Ns;
        if (Ns > 1) {
            const int step = N / (Ns * 8);
#pragma unroll
            for (int r = 1; r < 8; r++) v[q][r] = cmul(v[q][r], tw_get(tw, k * r * step, sign));
        }
        // 4-point DFTs of the even (0, 2, 4, 6) and odd (1, 3, 5, 7) points
        const C e0 = cadd(v[q][0], v[q][4]), e1 = csub(v[q][0], v[q][4]);
        const C e2 = cadd(v[q][2], v[q][6]), e3 = si(csub(v[q][2], v[q][6]));
        const C E0 = cadd(e0, e2), E1 = cadd(e1, e3), E2 = csub(e0, e2), E3 = csub(e1, e3);
        const C o0 = cadd(v[q][1], v[q][5]), o1 = csub(v[q][1], v[q][5]);
        const C o2 = cadd(v[q][3], v[q][7]), o3 = si(csub(v[q][3], v[q][7]));
        C O0 = cadd(o0, o2), O1 = cadd(o1, o3), O2 = csub(o0, o2), O3 = csub(o1, o3);
        // times W8^m = exp(s i pi m / 4), m = 1, 2, 3
        O1 = crs(h, cadd(O1, si(O1)));
        O2 = si(O2);
        O3 = crs(h, csub(si(O3), O3));
        C* dst = buf + line * N + (j / Ns) * Ns * 8 + k;
        dst[0] = cadd(E0, O0);
        dst[Ns] = cadd(E1, O1);
        dst[2 * Ns] = cadd(E2, O2);
        dst[3 * Ns] = cadd(E3, O3);
        dst[4 * Ns] = csub(E0, O0);
        dst[5 * Ns] = csub(E1, O1);
        dst[6 * Ns] = csub(E2, O2);
        dst[7 * Ns] = csub(E3, O3);
    }
    __syncthreads();
}

// In-place transform of the TOT/N lines held in buf (N a power of two).
template <int EPT, typename R, int TPB = FFT_THREADS>
__device__ __forceinline__ void fft_lds(cplx_t<R>* buf, int N, int log2N, int sign, const cplx_t<R>* __restrict__ tw) {
    int Ns = 1;
    if constexpr (EPT % 8 == 0) {
        // radix-8 passes while three or more factors of two remain, then one
        // radix-4 or radix-2 pass (the Stockham passes take the radices in any order)
        int rem = log2N;
        for (; rem >= 3; rem -= 3) {
            fft_pass_r8<EPT, R, TPB>(buf, N, Ns, sign, tw);
            Ns *= 8;
        }
        if (rem == 2) fft_pass_r4<EPT, R, TPB>(buf, N, Ns, sign, tw);
        else if (rem == 1) fft_pass_r2<EPT, R, TPB>(buf, N, Ns, sign, tw);
        return;
    }
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
