// fb_kernels.hpp -- argument blocks and device-side composite operations of
// the Fourier-Bessel transforms and FB-space operators (implemented in fb.hip).
#pragma once

#include <vector>

#include "common.hpp"

namespace chimera {

// Batched line FFT.  Line l = r + n_inner*q reads in[r*in_r + q*in_q + x] and
// writes out[r*out_r + q*out_q + x], x < N.
struct FftArgs {
    const double2* in;
    double2* out;
    int64_t nlines;
    int64_t n_inner;
    int64_t in_r, in_q, out_r, out_q;
    const double2* pre;    // per-x multiplier before the transform (nullable)
    const double2* post;   // per-x multiplier after the transform (nullable)
    const double* scale;   // real factor scale[x + N*r] after the transform (nullable)
    const double* filtr;   // absorbing profile: mask first nxfilt points, transform back (nullable)
    int nxfilt;
    int ghost;             // 0: none; 1: zero the row below r = 0; 2: eb_correction (mode scale, +-row 1)
    int nmodes;            // mode = q % nmodes (ghost != 0)
    int sign;              // -1 FFTW_FORWARD, +1 FFTW_BACKWARD
    int N, log2N;
    const double2* tw;
};

struct FbSeg {
    const double2* a0;
    const double2* a1;
    double2 c0, c1;
    const double* D;
    int64_t ldd;
    int refl;
    int acc;
};

struct FbOut {
    double2* p;
    double2 alpha0, alpha1;
    const double2* diag;
    double2 gamma;
    const double* scale;           // real factor per (x, k) after combination (nullable)
    int poiss;                     // Poisson-correction epilogue (maxwell_solvers.f90:131-164)
    const double2* poiss_np1;
    const double2* poiss_n;
    const double* poiss_w2;
};

struct FbJob {
    FbSeg seg[2];
    FbOut out[3];
    int nseg, nout, K, Ncol, beta;
};

struct FbGemmArgs {
    int nx;
    int64_t lda, ldo;
    const double* kx;
    double dt_inv;
};

enum { FB_ROT = 0, FB_GRAD = 1, FB_DIV = 2, FB_GRADDIV = 3 };

int32_t launch_fft(FftArgs a);
int32_t make_phase(double2* ph, const double* kx, double leftX, double s, int n);
int32_t launch_fb_gemm(const std::vector<FbJob>& jobs, const FbGemmArgs& g, int family);

// Deposition ghost fold in place (grid_deps.f90:80-85): row 1 -= row 0, row 0
// = 0 on each of nlines (nx1, nr1) planes.  Implemented in particles.hip.
int32_t ghost_fold(double2* a, int nx1, int nr1, int64_t nlines);

// Forward FB transform (fb_io.f90:18-98): in(nkx,0:nr,nmodes,ncomp) ->
// out(nkx,nkr,nmodes,ncomp), optionally times scale_tab(nkx,nkr,nmodes); with
// fold, the deposition ghost fold is first applied to in, in place.
int32_t fb_forward(double2* out, double2* in, double leftX, const double* kx_dev, const double* In_dev, int nkx,
                   int nr, int nmodes, int nkr, int ncomp, const double* scale_tab, int fold);
// Inverse FB transform of E (first 3 components of e_fb) and B into
// eb(nkx,0:nr,nmodes,6) (fb_io.f90:182-228); ghost as in FftArgs.
int32_t fb_inverse(double2* eb, const double2* e_fb, const double2* b_fb, int64_t b_comp_stride, double leftX,
                   const double* kx_dev, const double* Out_dev, int nkx, int nr, int nmodes, int nkr, int ghost);
// Absorbing layer (fb_io.f90:230-270, modefilt 0) on ncomp components in place.
int32_t fb_filter(double2* vec, double leftX, const double* kx_dev, const double* filtr_dev, int nxfilt, int nkx,
                  int nkr, int nmodes, int ncomp);
// FB-space operator (fb_math.f90), out = op(in); scale_tab (rot only) multiplies the result.
int32_t fb_op(int op, double2* out, const double2* in, const double* Dp, const double* Dm, const double* kx_dev,
              int nkx, int nkr, int nkO, int nkr_loc, const double* scale_tab);

}  // namespace chimera
