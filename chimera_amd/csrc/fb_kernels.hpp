// fb_kernels.hpp -- argument blocks and device-side composite operations of
// the Fourier-Bessel transforms and FB-space operators (implemented in fb.hip).
#pragma once

#include <vector>

#include "common.hpp"

namespace chimera {

// Batched line FFT on complex R (double: complex128, float: complex64).  Line
// l = r + n_inner*q reads in[r*in_r + q*in_q + x] and writes
// out[r*out_r + q*out_q + x], x < N.
template <typename R>
struct FftArgsT {
    typedef cplx_t<R> C;
    const C* in;
    C* out;
    int64_t nlines;
    int64_t n_inner;
    int64_t in_r, in_q, out_r, out_q;
    const C* pre;          // per-x multiplier before the transform (nullable)
    const C* post;         // per-x multiplier after the transform (nullable)
    const R* scale;        // real factor scale[x + N*r] after the transform (nullable)
    const R* filtr;        // absorbing profile: mask first nxfilt points, transform back (nullable)
    int nxfilt;
    int ghost;             // 0: none; 1: zero the row below r = 0; 2: eb_correction (mode scale, +-row 1);
                           // 3, 4: eb_correction_env (1/pi, +row 1 / -row 1)
    int nmodes;            // mode = q % nmodes (ghost != 0)
    int sign;              // -1 FFTW_FORWARD, +1 FFTW_BACKWARD
    int N, log2N;
    const C* tw;
    // per-x phase exp(ph_s*i*kx*leftX) (leftX from ph_leftX_dev when set) as
    // the pre (ph_at = 1) or post (ph_at = 2) multiplier: formed inside the
    // kernel for lines of up to 2048 points when its LDS has room, else by
    // k_phase into ph_table
    int ph_at;
    int tw_lds;            // set by launch_fft: twiddle table staged in LDS (EPT = 8)
    // set by launch_fft when N is not a power of two: the line is transformed
    // as a circular convolution of power-of-two length L >= 2N-1 (Bluestein),
    // bs_chirp/bs_ghat for the transform of sign `sign`, bs_chirp2/bs_ghat2 for
    // the transform back of fb_filtr (sign -sign)
    int L, log2L;
    const C* bs_chirp;
    const C* bs_ghat;
    const C* bs_chirp2;
    const C* bs_ghat2;
    double ph_s, ph_leftX;
    const double* ph_kx;
    const double* ph_leftX_dev;
    C* ph_table;
};
typedef FftArgsT<double> FftArgs;

// One K-segment of an FB GEMM job (see fb.hip): A = c0*T(a0) + c1*T(a1), times D.
template <typename R>
struct FbSegT {
    typedef cplx_t<R> C;
    const C* a0;
    const C* a1;
    double2 c0, c1;
    const R* D;
    int64_t ldd;
    int refl;
    int acc;
};

template <typename R>
struct FbOutT {
    typedef cplx_t<R> C;
    C* p;
    double2 alpha0, alpha1;
    const C* diag;
    double2 gamma;
    const R* scale;                // real factor per (x, k) after combination (nullable)
    int poiss;                     // Poisson-correction epilogue (maxwell_solvers.f90:131-164)
    const C* poiss_np1;
    const C* poiss_n;
    const R* poiss_w2;
};

template <typename R>
struct FbJobT {
    FbSegT<R> seg[2];
    FbOutT<R> out[3];
    int nseg, nout, K, Ncol, beta;
};
typedef FbSegT<double> FbSeg;
typedef FbOutT<double> FbOut;
typedef FbJobT<double> FbJob;

struct FbGemmArgs {
    int nx;
    int64_t lda, ldo;
    const double* kx;
    double dt_inv;
    int buf_ok;            // set by launch_fb_gemm: every operand offset fits the 31 bits of a buffer load
};

enum { FB_ROT = 0, FB_GRAD = 1, FB_DIV = 2, FB_GRADDIV = 3 };
// form of the complex64 GEMMs (chimera_set_mixed_gemm)
enum { MIXED_GEMM_BF16X6 = 0, MIXED_GEMM_FP32 = 1 };

template <typename R> int32_t launch_fft(FftArgsT<R> a);
// EUNSUP unless the x transform handles length N
int32_t fft_length_check(int N);
// ph[i] = exp(s*i*kx[i]*leftX), leftX read from leftX_dev on the device when that is set.
template <typename R>
int32_t make_phase(cplx_t<R>* ph, const double* kx, double leftX, double s, int n, const double* leftX_dev = nullptr);
template <typename R> int32_t launch_fb_gemm(const std::vector<FbJobT<R>>& jobs, const FbGemmArgs& g, int family);

// Deposition ghost fold in place (grid_deps.f90:80-85): row 1 -= row 0, row 0
// = 0 on each of nlines (nx1, nr1) planes.  Implemented in particles.hip.
template <typename R> int32_t ghost_fold(cplx_t<R>* a, int nx1, int nr1, int64_t nlines);

// Forward FB transform (fb_io.f90:18-98): in(nkx,0:nr,nmodes,ncomp) ->
// out(nkx,nkr,nmodes,ncomp), optionally times scale_tab(nkx,nkr,nmodes); with
// fold, the deposition ghost fold is first applied to in, in place.  A
// non-null leftX_dev replaces leftX by the device value at run time.
template <typename R>
int32_t fb_forward_t(cplx_t<R>* out, cplx_t<R>* in, double leftX, const double* kx_dev, const R* In_dev, int nkx,
                     int nr, int nmodes, int nkr, int ncomp, const R* scale_tab, int fold,
                     const double* leftX_dev = nullptr);
// Inverse FB transform of E (first 3 components of e_fb) and B into
// eb(nkx,0:nr,nmodes,6) (fb_io.f90:182-228); ghost as in FftArgs.
template <typename R>
int32_t fb_inverse_t(cplx_t<R>* eb, const cplx_t<R>* e_fb, const cplx_t<R>* b_fb, int64_t b_comp_stride, double leftX,
                     const double* kx_dev, const R* Out_dev, int nkx, int nr, int nmodes, int nkr, int ghost,
                     const double* leftX_dev = nullptr);
// Absorbing layer (fb_io.f90:230-270, modefilt 0) on ncomp components in place.
template <typename R>
int32_t fb_filter_t(cplx_t<R>* vec, double leftX, const double* kx_dev, const R* filtr_dev, int nxfilt, int nkx,
                    int nkr, int nmodes, int ncomp);
// FB-space operator (fb_math.f90), out = op(in); scale_tab (rot only) multiplies the result.
template <typename R>
int32_t fb_op_t(int op, cplx_t<R>* out, const cplx_t<R>* in, const R* Dp, const R* Dm, const double* kx_dev, int nkx,
                int nkr, int nkO, int nkr_loc, const R* scale_tab);
// j += (graddiv j + (gp - gn)*dt_inv)*w2_inv, one Poisson-correction iteration.
template <typename R>
int32_t fb_graddiv_poisson_t(cplx_t<R>* j, const cplx_t<R>* gn, const cplx_t<R>* gp, double dt_inv, const R* w2,
                             const R* Dp, const R* Dm, const double* kx_dev, int nkx, int nkr, int nkO);

// FP64 names used by the drop-in entry points
inline int32_t fb_forward(double2* out, double2* in, double leftX, const double* kx_dev, const double* In_dev, int nkx,
                          int nr, int nmodes, int nkr, int ncomp, const double* scale_tab, int fold) {
    return fb_forward_t<double>(out, in, leftX, kx_dev, In_dev, nkx, nr, nmodes, nkr, ncomp, scale_tab, fold);
}
inline int32_t fb_inverse(double2* eb, const double2* e_fb, const double2* b_fb, int64_t b_comp_stride, double leftX,
                          const double* kx_dev, const double* Out_dev, int nkx, int nr, int nmodes, int nkr, int ghost) {
    return fb_inverse_t<double>(eb, e_fb, b_fb, b_comp_stride, leftX, kx_dev, Out_dev, nkx, nr, nmodes, nkr, ghost);
}
inline int32_t fb_filter(double2* vec, double leftX, const double* kx_dev, const double* filtr_dev, int nxfilt,
                         int nkx, int nkr, int nmodes, int ncomp) {
    return fb_filter_t<double>(vec, leftX, kx_dev, filtr_dev, nxfilt, nkx, nkr, nmodes, ncomp);
}
inline int32_t fb_op(int op, double2* out, const double2* in, const double* Dp, const double* Dm, const double* kx_dev,
                     int nkx, int nkr, int nkO, int nkr_loc, const double* scale_tab) {
    return fb_op_t<double>(op, out, in, Dp, Dm, kx_dev, nkx, nkr, nkO, nkr_loc, scale_tab);
}
inline int32_t ghost_fold_d(double2* a, int nx1, int nr1, int64_t nlines) {
    return ghost_fold<double>(a, nx1, nr1, nlines);
}

}  // namespace chimera
