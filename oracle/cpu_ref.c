/*
 * cpu_ref.c -- CPU FP64 restatement of CHIMERA's particle-in-cell step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: the parity tests, the
 * smoke check and the CPU-baseline leg of the benchmark call it as the
 * checker.  The product (chimera_amd, libchimera_hip) never links or calls it.
 *
 * Every routine restates one subroutine of the reference's Fortran (cited as
 * file:line, relative to the reference repository) with the same loop order,
 * the same operation order inside each expression and the same array layout:
 * Fortran column-major arrays, complex(kind=8) as interleaved (re, im) doubles,
 * dimension arguments with the meaning they have in the Fortran declaration
 * (e.g. `nx` is the last x index of a 0:nx array).  Scalars are passed by
 * value.  OpenMP pragmas sit where the reference has them, so the file doubles
 * as the timed CPU baseline; the serial loops (deposition) stay serial.
 *
 * The FFTs of fb_io.f90 come from FFTW3 in the reference; FFTW is not part of
 * this code base, so `fft_line` below is a textbook radix-2 / direct DFT with
 * FFTW's sign and scaling conventions (unnormalised, FORWARD = exp(-2 pi i jk/N)).
 *
 * Compile without -ffast-math and with -ffp-contract=off so that results are
 * plain IEEE double arithmetic in the order written.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cplx;

static inline cplx cmk(double re, double im) { cplx z = {re, im}; return z; }
static inline cplx cmul(cplx a, cplx b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cplx cadd(cplx a, cplx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline cplx csub(cplx a, cplx b) { return cmk(a.re - b.re, a.im - b.im); }
static inline cplx crs(double s, cplx a) { return cmk(s * a.re, s * a.im); }      /* real * complex */
static inline cplx cmuli(cplx a) { return cmk(-a.im, a.re); }                     /* i * a */
static inline cplx cconj(cplx a) { return cmk(a.re, -a.im); }
static inline cplx cneg(cplx a) { return cmk(-a.re, -a.im); }

#define PI_D 3.14159265358979323846

/* ------------------------------------------------------------------------- */
/* particle_tools.f90                                                        */
/* ------------------------------------------------------------------------- */

/* Boris pusher.  particle_tools.f90:18-56. momenta(3,np) in/out, Fld(6,np). */
void ref_push_velocs(double *momenta, const double *Fld, double dt, int32_t np)
{
    const double dt_2 = 0.5 * dt;
#pragma omp parallel for schedule(static)
    for (int64_t ip = 0; ip < np; ip++) {
        const double *F = Fld + 6 * ip;
        double *u = momenta + 3 * ip;
        double Um[3], t[3], s[3], U0[3], Up[3];
        for (int c = 0; c < 3; c++) Um[c] = u[c] + dt_2 * F[c];
        const double gamma = sqrt(1.0 + (Um[0] * Um[0] + Um[1] * Um[1] + Um[2] * Um[2]));
        for (int c = 0; c < 3; c++) t[c] = dt_2 * F[3 + c] / gamma;
        const double t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
        for (int c = 0; c < 3; c++) s[c] = 2 * t[c] / (1 + t2);
        U0[0] = Um[0] + Um[1] * t[2] - Um[2] * t[1];
        U0[1] = Um[1] - Um[0] * t[2] + Um[2] * t[0];
        U0[2] = Um[2] + Um[0] * t[1] - Um[1] * t[0];
        Up[0] = Um[0] + U0[1] * s[2] - U0[2] * s[1];
        Up[1] = Um[1] - U0[0] * s[2] + U0[2] * s[0];
        Up[2] = Um[2] + U0[0] * s[1] - U0[1] * s[0];
        for (int c = 0; c < 3; c++) u[c] = Up[c] + dt_2 * F[c];
    }
}

/* Leap-frog position push.  particle_tools.f90:58-82. */
void ref_push_coords(double *coord, const double *momenta, double *coord_cntr, double dt, int32_t np)
{
#pragma omp parallel for schedule(static)
    for (int64_t ip = 0; ip < np; ip++) {
        const double *u = momenta + 3 * ip;
        double *x = coord + 3 * ip;
        double *xc = coord_cntr + 3 * ip;
        const double dt_gp = dt / sqrt(1.0 + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
        for (int c = 0; c < 3; c++) {
            const double x0 = x[c];
            const double x1 = x0 + u[c] * dt_gp;
            x[c] = x1;
            xc[c] = 0.5 * (x0 + x1);
        }
    }
}

/* FixedCell lattice generator.  particle_tools.f90:84-128.
 * coord(4,np) out (x, y, z, r), *indPart = number written.  RandPackO(nx,nr),
 * PackO complex(PPC). */
void ref_genparts(double *coord, int32_t *indPart, const double *Xgrid, const double *Rgrid,
                  const double *RandPackO, const double *PackX, const double *PackR,
                  const double *PackO, int32_t np, int32_t nx, int32_t nr, int32_t PPC)
{
    memset(coord, 0, sizeof(double) * 4 * (size_t)np);
    const double pi = 4.0 * atan(1.0);
    const double dr_2 = 0.5 * (Rgrid[1] - Rgrid[0]);
    int32_t ind = 0;
    for (int ir = 0; ir < nr - 1; ir++) {
        for (int ix = 0; ix < nx - 1; ix++) {
            const double r0 = Rgrid[ir] + dr_2, r1 = Rgrid[ir + 1] + dr_2;
            const double x0 = Xgrid[ix], x1 = Xgrid[ix + 1];
            const double ang = 2.0 * pi * RandPackO[ix + (int64_t)nx * ir];
            const cplx Oshft = cmk(cos(ang), sin(ang));
            for (int ip = 0; ip < PPC; ip++) {
                const double xc = x0 + (x1 - x0) * PackX[ip];
                const double rc = r0 + (r1 - r0) * PackR[ip];
                if (rc <= 0) continue;
                const cplx O = cmul(cmk(PackO[2 * ip], PackO[2 * ip + 1]), Oshft);
                coord[4 * (int64_t)ind + 0] = xc;
                coord[4 * (int64_t)ind + 1] = rc * O.im;
                coord[4 * (int64_t)ind + 2] = rc * O.re;
                coord[4 * (int64_t)ind + 3] = rc;
                ind++;
            }
        }
    }
    *indPart = ind;
}

/* Domain filter.  particle_tools.f90:130-153.  indx2stay zero-based. */
void ref_sortpartsout(int32_t *indx2stay, int32_t *num2stay, const double *coord, const double *lims, int32_t np)
{
    int32_t n = 0;
    memset(indx2stay, 0, sizeof(int32_t) * (size_t)np);
    for (int64_t ip = 0; ip < np; ip++) {
        const double x = coord[3 * ip];
        const double r2 = coord[3 * ip + 1] * coord[3 * ip + 1] + coord[3 * ip + 2] * coord[3 * ip + 2];
        if (x >= lims[0] && x <= lims[1] && r2 >= lims[2] && r2 <= lims[3]) indx2stay[n++] = (int32_t)ip;
    }
    *num2stay = n;
}

/* Chunk binning.  particle_tools.f90:155-208.  Xgrid(0:nx). */
void ref_chunk_coords_boundaries(int8_t *chunked_indx, int32_t *IndInChnk, int32_t *GoOut,
                                 const double *coord, const double *lims, const double *Xgrid,
                                 int32_t nchnk, int32_t np, int32_t nx)
{
    int64_t *num = (int64_t *)calloc((size_t)nchnk, sizeof(int64_t));
    double inv;
    if (nchnk > 1) inv = 1. / (Xgrid[(nx + 1) / nchnk] - Xgrid[0]);
    else inv = 1. / (Xgrid[nx] - Xgrid[0]);
    int32_t out = 0;
    for (int64_t ip = 0; ip < np; ip++) {
        chunked_indx[ip] = -2;
        const int ichnk = (int)floor((coord[3 * ip] - Xgrid[0]) * inv);
        const double x = coord[3 * ip];
        const double r2 = coord[3 * ip + 1] * coord[3 * ip + 1] + coord[3 * ip + 2] * coord[3 * ip + 2];
        if (x >= lims[0] && x <= lims[1] && r2 >= lims[2] && r2 <= lims[3]) {
            if (ichnk >= 0 && ichnk < nchnk) num[ichnk]++;
            chunked_indx[ip] = (int8_t)ichnk;
        } else {
            out++;
        }
    }
    IndInChnk[0] = 0;
    for (int c = 0; c < nchnk; c++) IndInChnk[c + 1] = IndInChnk[c] + (int32_t)num[c];
    *GoOut = out;
    free(num);
}

/* Permutation gather.  particle_tools.f90:270-296 (vec) and :298-324 (scl).
 * idx is zero-based (as produced by numpy argsort), length n; dat has ncomp
 * rows and at least n columns. */
void ref_align_data(double *dat, const int64_t *idx, int32_t ncomp, int32_t n, int32_t np0)
{
    (void)np0;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)ncomp * (size_t)(n > 0 ? n : 1));
#pragma omp parallel for
    for (int64_t ip = 0; ip < n; ip++)
        for (int c = 0; c < ncomp; c++) tmp[ncomp * ip + c] = dat[ncomp * idx[ip] + c];
#pragma omp parallel for
    for (int64_t ip = 0; ip < n; ip++)
        for (int c = 0; c < ncomp; c++) dat[ncomp * ip + c] = tmp[ncomp * ip + c];
    free(tmp);
}

/* ------------------------------------------------------------------------- */
/* grid_deps.f90                                                             */
/* ------------------------------------------------------------------------- */

/* Shared per-particle geometry of dep_curr / dep_dens / proj_fld.
 * grid_deps.f90:46-66.  sgn = -1 for deposition ((y - i z)/r), +1 for gather. */
typedef struct { int64_t ix, ir; double S0[2][2]; cplx phase[16]; } pgeom;

/* Cell, linear weights and azimuthal phases of one particle (grid_deps.f90:46-59).
 * Returns 0 when the cell lies outside the arrays: the reference indexes
 * without an x bound there (grid_deps.f90:46-47) and writes or reads outside
 * its array; this restatement skips such particles, as the GPU kernels do. */
static inline int ref_geom(pgeom *g, double xp, double yp, double zp, double rp, double leftX,
                           const double *Rgrid, double dx_inv, double dr_inv, int nx, int nr, int nkO, double sgn)
{
    const double fx = (xp - leftX) * dx_inv, fr = (rp - Rgrid[0]) * dr_inv;
    if (!(fx >= 0.0 && fx < (double)nx) || !(fr >= 0.0 && fr < (double)nr)) return 0;
    g->ix = (int64_t)floor(fx);
    g->ir = (int64_t)floor(fr);
    g->S0[1][0] = (xp - leftX) * dx_inv - (double)g->ix;
    g->S0[0][0] = 1.0 - g->S0[1][0];
    g->S0[1][1] = (rp - Rgrid[g->ir]) * dr_inv;
    g->S0[0][1] = 1.0 - g->S0[1][1];
    cplx ph;
    if (rp > 0.0) ph = cmk(yp / rp, sgn * zp / rp);
    else ph = cmk(0.0, 0.0);
    g->phase[0] = cmk(1.0, 0.0);
    for (int iO = 1; iO <= nkO; iO++) g->phase[iO] = cmul(g->phase[iO - 1], ph);
    return 1;
}

/* Current deposition, serial.  grid_deps.f90:18-87.
 * curr(0:nx,0:nr,0:nkO,3) complex, accumulated, then ghost fold. */
void ref_dep_curr(const double *coord, const double *momenta, const double *wghts, double *curr_,
                  double leftX, const double *Rgrid, double dx_inv, double dr_inv,
                  int32_t np, int32_t nx, int32_t nr, int32_t nkO)
{
    cplx *curr = (cplx *)curr_;
    const int64_t sx = 1, sr = nx + 1, sm = (int64_t)(nx + 1) * (nr + 1), sl = sm * (nkO + 1);
    for (int64_t ip = 0; ip < np; ip++) {
        const double wp = wghts[ip];
        if (wp == 0.0) continue;
        const double xp = coord[3 * ip], yp = coord[3 * ip + 1], zp = coord[3 * ip + 2];
        double v[3] = {momenta[3 * ip], momenta[3 * ip + 1], momenta[3 * ip + 2]};
        const double rp = sqrt(yp * yp + zp * zp);
        if (rp >= Rgrid[nr] || (fabs(v[0]) + fabs(v[1]) + fabs(v[2])) == 0.0) continue;
        const double gp = sqrt(1 + v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int c = 0; c < 3; c++) v[c] = v[c] * wp / gp;
        pgeom g;
        if (!ref_geom(&g, xp, yp, zp, rp, leftX, Rgrid, dx_inv, dr_inv, nx, nr, nkO, -1.0)) continue;
        double cp[2][2];
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) cp[a][b] = g.S0[a][0] * g.S0[b][1];
        for (int l = 0; l < 3; l++)
            for (int iO = 0; iO <= nkO; iO++) {
                const cplx pv = crs(v[l], g.phase[iO]);
                for (int b = 0; b < 2; b++)
                    for (int a = 0; a < 2; a++) {
                        cplx *dst = curr + (g.ix + a) * sx + (g.ir + b) * sr + iO * sm + l * sl;
                        *dst = cadd(*dst, crs(cp[a][b], pv));
                    }
            }
    }
    for (int l = 0; l < 3; l++)
        for (int iO = 0; iO <= nkO; iO++)
            for (int64_t ix = 0; ix <= nx; ix++) {
                cplx *r0 = curr + ix + iO * sm + l * sl;
                r0[sr] = csub(r0[sr], r0[0]);
                r0[0] = cmk(0.0, 0.0);
            }
}

/* Charge deposition, serial.  grid_deps.f90:89-147.  dens(0:nx,0:nr,0:nkO). */
void ref_dep_dens(const double *coord, const double *wghts, double *dens_, double leftX,
                  const double *Rgrid, double dx_inv, double dr_inv,
                  int32_t np, int32_t nx, int32_t nr, int32_t nkO)
{
    cplx *dens = (cplx *)dens_;
    const int64_t sr = nx + 1, sm = (int64_t)(nx + 1) * (nr + 1);
    for (int64_t ip = 0; ip < np; ip++) {
        const double wp = wghts[ip];
        if (wp == 0.0) continue;
        const double xp = coord[3 * ip], yp = coord[3 * ip + 1], zp = coord[3 * ip + 2];
        const double rp = sqrt(yp * yp + zp * zp);
        if (rp >= Rgrid[nr]) continue;
        pgeom g;
        if (!ref_geom(&g, xp, yp, zp, rp, leftX, Rgrid, dx_inv, dr_inv, nx, nr, nkO, -1.0)) continue;
        double dp[2][2];
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) dp[a][b] = g.S0[a][0] * g.S0[b][1] * wp;
        for (int iO = 0; iO <= nkO; iO++)
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) {
                    cplx *dst = dens + (g.ix + a) + (g.ir + b) * sr + iO * sm;
                    *dst = cadd(*dst, crs(dp[a][b], g.phase[iO]));
                }
    }
    for (int iO = 0; iO <= nkO; iO++)
        for (int64_t ix = 0; ix <= nx; ix++) {
            cplx *r0 = dens + ix + iO * sm;
            r0[sr] = csub(r0[sr], r0[0]);
            r0[0] = cmk(0.0, 0.0);
        }
}

/* Field gather.  grid_deps.f90:149-217.  Fld(0:nx,0:nr,0:nkO,6) complex,
 * Fld_tot(6,np) accumulated. */
void ref_proj_fld(const double *coord, const double *wghts, const double *Fld_, double *Fld_tot,
                  double leftX, const double *Rgrid, double dx_inv, double dr_inv,
                  int32_t np, int32_t nx, int32_t nr, int32_t nkO)
{
    const cplx *Fld = (const cplx *)Fld_;
    const int64_t sr = nx + 1, sm = (int64_t)(nx + 1) * (nr + 1), sl = sm * (nkO + 1);
#pragma omp parallel for schedule(static)
    for (int64_t ip = 0; ip < np; ip++) {
        const double wp = wghts[ip];
        if (wp == 0.0) continue;
        const double xp = coord[3 * ip], yp = coord[3 * ip + 1], zp = coord[3 * ip + 2];
        const double rp = sqrt(yp * yp + zp * zp);
        if (rp >= Rgrid[nr]) continue;
        pgeom g;
        if (!ref_geom(&g, xp, yp, zp, rp, leftX, Rgrid, dx_inv, dr_inv, nx, nr, nkO, +1.0)) continue;
        cplx pc[16][2][2];
        for (int iO = 0; iO <= nkO; iO++)
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) pc[iO][a][b] = crs(g.S0[b][1] * g.S0[a][0], g.phase[iO]);
        for (int l = 0; l < 6; l++) {
            double acc = 0.0;
            for (int iO = 0; iO <= nkO; iO++)
                for (int b = 0; b < 2; b++)
                    for (int a = 0; a < 2; a++) {
                        const cplx f = Fld[(g.ix + a) + (g.ir + b) * sr + iO * sm + l * sl];
                        acc += pc[iO][a][b].re * f.re - pc[iO][a][b].im * f.im;
                    }
            Fld_tot[6 * ip + l] = Fld_tot[6 * ip + l] + acc;
        }
    }
}

/* Mode normalisation and ghost row.  grid_deps.f90:219-266.
 * eb_spc(0:nx,0:nr,nkO,6): here nkO is the NUMBER of modes. */
void ref_eb_correction(double *eb_, int32_t nx, int32_t nr, int32_t nkO)
{
    cplx *eb = (cplx *)eb_;
    const double pi = 4.0 * atan(1.0);
    const double pi_inv = 1. / pi, pi2_inv = 0.5 * pi_inv;
    const int64_t sr = nx + 1, sm = (int64_t)(nx + 1) * (nr + 1), sl = sm * nkO;
    for (int l = 0; l < 6; l++)
        for (int iO = 0; iO < nkO; iO++) {
            const double f = (iO == 0) ? pi2_inv : pi_inv;
            for (int64_t j = 0; j < sm; j++) eb[j + iO * sm + l * sl] = crs(f, eb[j + iO * sm + l * sl]);
        }
    for (int l = 0; l < 6; l++)
        for (int iO = 0; iO < nkO; iO++)
            for (int64_t ix = 0; ix <= nx; ix++) {
                cplx *c = eb + ix + iO * sm + l * sl;
                c[0] = (iO == 0) ? c[sr] : cneg(c[sr]);
            }
}

/* ------------------------------------------------------------------------- */
/* FFT with FFTW's conventions (the reference links FFTW3, fb_io.f90:37-38)  */
/* ------------------------------------------------------------------------- */

/* In-place length-n DFT of a, sign -1 (FFTW_FORWARD) or +1 (FFTW_BACKWARD),
 * unnormalised.  Radix-2 decimation in time for powers of two, direct O(n^2)
 * sum in long double otherwise.  Twiddles from cos/sin of the exact angle. */
static void fft_line(cplx *a, int64_t n, int sign, cplx *work)
{
    if (n <= 1) return;
    if ((n & (n - 1)) == 0) {
        for (int64_t i = 1, j = 0; i < n; i++) {
            int64_t bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { cplx t = a[i]; a[i] = a[j]; a[j] = t; }
        }
        for (int64_t len = 2; len <= n; len <<= 1) {
            const int64_t half = len >> 1, step = n / len;
            for (int64_t k = 0; k < half; k++) {
                const double ang = sign * 2.0 * PI_D * (double)(k * step) / (double)n;
                const cplx w = cmk(cos(ang), sin(ang));
                for (int64_t i = k; i < n; i += len) {
                    const cplx u = a[i], v = cmul(a[i + half], w);
                    a[i] = cadd(u, v);
                    a[i + half] = csub(u, v);
                }
            }
        }
    } else {
        /* direct sum in long double (64-bit significand) with a long double
         * twiddle table: the result is the DFT rounded once to double, more
         * accurate than any FFT, so that the oracle's own rounding stays well
         * below the differences a parity test measures */
        long double *twr = (long double *)malloc(sizeof(long double) * 2 * (size_t)n);
        long double *twi = twr + n;
        for (int64_t k = 0; k < n; k++) {
            const long double ang = (long double)sign * 2.0L * 3.141592653589793238462643383279502884L *
                                    (long double)k / (long double)n;
            twr[k] = cosl(ang);
            twi[k] = sinl(ang);
        }
        for (int64_t k = 0; k < n; k++) {
            long double sr = 0, si = 0;
            int64_t jk = 0;
            for (int64_t j = 0; j < n; j++) {
                const long double c = twr[jk], s = twi[jk];
                sr += (long double)a[j].re * c - (long double)a[j].im * s;
                si += (long double)a[j].re * s + (long double)a[j].im * c;
                jk += k;
                if (jk >= n) jk -= n;
            }
            work[k] = cmk((double)sr, (double)si);
        }
        free(twr);
        memcpy(a, work, sizeof(cplx) * (size_t)n);
    }
}

/* ------------------------------------------------------------------------- */
/* fb_io.f90                                                                  */
/* ------------------------------------------------------------------------- */

/* Forward Fourier-Bessel transform of a vector field.  fb_io.f90:18-59.
 * vec(nkx,0:nr,nkO,ncomp) -> vec_fb(nkx,nkr,nkO,ncomp); In(nr,nkr,nkO);
 * nkO is the number of modes.  ncomp = 3 restates fb_vec_in, ncomp = 1
 * restates fb_scl_in (fb_io.f90:61-98). */
void ref_fb_in(double *vec_fb_, const double *vec_, double leftX, const double *kx, const double *In,
               int32_t nkx, int32_t nr, int32_t nkO, int32_t nkr, int32_t ncomp)
{
    cplx *vec_fb = (cplx *)vec_fb_;
    const cplx *vec = (const cplx *)vec_;
    memset(vec_fb, 0, sizeof(cplx) * (size_t)nkx * nkr * nkO * ncomp);
    cplx *shiftX = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
    for (int64_t i = 0; i < nkx; i++) shiftX[i] = cmk(cos(leftX * kx[i]), -sin(leftX * kx[i]));
    for (int l = 0; l < ncomp; l++)
        for (int iO = 0; iO < nkO; iO++) {
#pragma omp parallel
            {
                cplx *A = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
                cplx *W = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
#pragma omp for schedule(static)
                for (int ik = 0; ik < nkr; ik++) {
                    for (int64_t i = 0; i < nkx; i++) A[i] = cmk(0.0, 0.0);
                    for (int ir = 1; ir <= nr; ir++) {
                        const double c = In[(ir - 1) + (int64_t)nr * ik + (int64_t)nr * nkr * iO];
                        const cplx *src = vec + (int64_t)nkx * (ir + (int64_t)(nr + 1) * (iO + (int64_t)nkO * l));
                        for (int64_t i = 0; i < nkx; i++) A[i] = cadd(A[i], crs(c, src[i]));
                    }
                    fft_line(A, nkx, -1, W);
                    cplx *dst = vec_fb + (int64_t)nkx * (ik + (int64_t)nkr * (iO + (int64_t)nkO * l));
                    for (int64_t i = 0; i < nkx; i++) dst[i] = cadd(dst[i], cmul(A[i], shiftX[i]));
                }
                free(A); free(W);
            }
        }
    free(shiftX);
}

/* Inverse Fourier-Bessel transform of E and B.  fb_io.f90:182-228.
 * e_fb(nkx,nkr,nkO,>=3) (only components 1..3 read), b_fb(nkx,nkr,nkO,3),
 * Out(nkr,nr,nkO) -> eb_spc(nkx,0:nr,nkO,6), zeroed first. */
void ref_fb_eb_out(double *eb_, const double *e_fb_, const double *b_fb_, double leftX, const double *kx,
                   const double *Out, int32_t nkx, int32_t nr, int32_t nkO, int32_t nkr)
{
    cplx *eb = (cplx *)eb_;
    const cplx *e_fb = (const cplx *)e_fb_, *b_fb = (const cplx *)b_fb_;
    memset(eb, 0, sizeof(cplx) * (size_t)nkx * (nr + 1) * nkO * 6);
    cplx *shiftX = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
    for (int64_t i = 0; i < nkx; i++) shiftX[i] = cmk(cos(leftX * kx[i]), sin(leftX * kx[i]));
    const int64_t fb_l = (int64_t)nkx * nkr * nkO;
    for (int l = 0; l < 3; l++)
        for (int iO = 0; iO < nkO; iO++) {
#pragma omp parallel
            {
                cplx *A = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
                cplx *W = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
#pragma omp for schedule(static)
                for (int ir = 1; ir <= nr; ir++) {
                    for (int half = 0; half < 2; half++) {
                        const cplx *src = half == 0 ? e_fb : b_fb;
                        for (int64_t i = 0; i < nkx; i++) A[i] = cmk(0.0, 0.0);
                        for (int ik = 0; ik < nkr; ik++) {
                            const double c = Out[ik + (int64_t)nkr * (ir - 1) + (int64_t)nkr * nr * iO];
                            const cplx *s = src + (int64_t)nkx * (ik + (int64_t)nkr * iO) + l * fb_l;
                            for (int64_t i = 0; i < nkx; i++) A[i] = cadd(A[i], crs(c, s[i]));
                        }
                        for (int64_t i = 0; i < nkx; i++) A[i] = cmul(A[i], shiftX[i]);
                        fft_line(A, nkx, +1, W);
                        cplx *dst = eb + (int64_t)nkx * (ir + (int64_t)(nr + 1) * (iO + (int64_t)nkO * (l + 3 * half)));
                        for (int64_t i = 0; i < nkx; i++) dst[i] = cadd(dst[i], A[i]);
                    }
                }
                free(A); free(W);
            }
        }
    free(shiftX);
}

/* Absorbing layer in x.  fb_io.f90:230-308, modefilt = 0 ('left') only:
 * the other two modes index filtr out of its bounds in the reference
 * (fb_io.f90:279, :296) and are never requested by the driver.
 * vec(nkx,nkr,nkO,3) in/out, filtr(nxfilt).  Returns -1 for modefilt != 0. */
int ref_fb_filtr(double *vec_, double leftX, const double *kx, const double *filtr, int32_t modefilt,
                 int32_t nkx, int32_t nkr, int32_t nkO, int32_t nxfilt)
{
    if (modefilt != 0) return -1;
    cplx *vec = (cplx *)vec_;
    cplx *shiftX = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
    cplx *shiftX_inv = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
    for (int64_t i = 0; i < nkx; i++) {
        shiftX[i] = cmk(cos(leftX * kx[i]), sin(leftX * kx[i]));
        /* 1/(nkx*shiftX): complex reciprocal of (N cos, N sin) */
        const double a = (double)nkx * shiftX[i].re, b = (double)nkx * shiftX[i].im;
        const double d = a * a + b * b;
        shiftX_inv[i] = cmk(a / d, -b / d);
    }
    for (int l = 0; l < 3; l++)
        for (int iO = 0; iO < nkO; iO++) {
#pragma omp parallel
            {
                cplx *A = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
                cplx *W = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
#pragma omp for schedule(static)
                for (int ik = 0; ik < nkr; ik++) {
                    cplx *v = vec + (int64_t)nkx * (ik + (int64_t)nkr * (iO + (int64_t)nkO * l));
                    for (int64_t i = 0; i < nkx; i++) A[i] = cmul(v[i], shiftX[i]);
                    fft_line(A, nkx, +1, W);
                    for (int64_t i = 0; i < nxfilt && i < nkx; i++) A[i] = crs(filtr[i], A[i]);
                    fft_line(A, nkx, -1, W);
                    for (int64_t i = 0; i < nkx; i++) v[i] = cmul(A[i], shiftX_inv[i]);
                }
                free(A); free(W);
            }
        }
    free(shiftX); free(shiftX_inv);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* fb_math.f90: FB-space vector calculus (nkO = max mode, arrays 0:nkO)      */
/* ------------------------------------------------------------------------- */

#define FB(a, x, k, m, c) (a)[(x) + (int64_t)nkx * ((k) + (int64_t)nkr * ((m) + (int64_t)(nmod) * (c)))]
#define DM(D, k, kl, m) (D)[(k) + (int64_t)nkr * ((kl) + (int64_t)nkr_loc * (m))]

/* ext(x) = -conj(v(-kx, m=1)): fb_math.f90:35-36. */
static inline int64_t refl(int64_t x, int64_t nkx) { return x == 0 ? 0 : nkx - x; }

/* Curl.  fb_math.f90:18-94.  vec_fb(nkx,nkr,0:nkO,3) -> vec_fb_loc(nkx,nkr_loc,0:nkO,3). */
void ref_fb_rot(double *loc_, const double *vec_, const double *DpS2S, const double *DmS2S, const double *kx,
                int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc)
{
    cplx *loc = (cplx *)loc_;
    const cplx *v = (const cplx *)vec_;
    const int nmod = nkO + 1;
    memset(loc, 0, sizeof(cplx) * (size_t)nkx * nkr_loc * nmod * 3);
    for (int iO = 0; iO <= nkO; iO++) {
#pragma omp parallel
        {
            cplx *sl = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
#pragma omp for schedule(static)
            for (int kl = 0; kl < nkr_loc; kl++) {
                cplx *o1 = &loc[(int64_t)nkx * (kl + (int64_t)nkr_loc * (iO + (int64_t)nmod * 0))];
                cplx *o2 = &loc[(int64_t)nkx * (kl + (int64_t)nkr_loc * (iO + (int64_t)nmod * 1))];
                cplx *o3 = &loc[(int64_t)nkx * (kl + (int64_t)nkr_loc * (iO + (int64_t)nmod * 2))];
                for (int64_t x = 0; x < nkx; x++) {
                    o2[x] = csub(o2[x], cmuli(crs(kx[x], FB(v, x, kl, iO, 2))));
                    o3[x] = cadd(o3[x], cmuli(crs(kx[x], FB(v, x, kl, iO, 1))));
                }
                if (iO < nkO) {
                    for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                    for (int k = 0; k < nkr; k++) {
                        const double d = DM(DpS2S, k, kl, iO);
                        for (int64_t x = 0; x < nkx; x++)
                            sl[x] = csub(sl[x], crs(d, csub(cmuli(FB(v, x, k, iO + 1, 1)), FB(v, x, k, iO + 1, 2))));
                    }
                    for (int64_t x = 0; x < nkx; x++) o1[x] = cadd(o1[x], sl[x]);
                    for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                    for (int k = 0; k < nkr; k++) {
                        const double d = DM(DpS2S, k, kl, iO);
                        for (int64_t x = 0; x < nkx; x++) sl[x] = cadd(sl[x], crs(d, FB(v, x, k, iO + 1, 0)));
                    }
                    for (int64_t x = 0; x < nkx; x++) {
                        o2[x] = cadd(o2[x], cmuli(sl[x]));
                        o3[x] = csub(o3[x], sl[x]);
                    }
                }
                /* m-1 neighbour; for iO == 0 the reflected conjugate of mode 1 */
                for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                for (int k = 0; k < nkr; k++) {
                    const double d = DM(DmS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++) {
                        cplx a2, a3;
                        if (iO > 0) { a2 = FB(v, x, k, iO - 1, 1); a3 = FB(v, x, k, iO - 1, 2); }
                        else if (nkO >= 1) {
                            a2 = cneg(cconj(FB(v, refl(x, nkx), k, 1, 1)));
                            a3 = cneg(cconj(FB(v, refl(x, nkx), k, 1, 2)));
                        } else { a2 = cmk(0, 0); a3 = cmk(0, 0); }
                        sl[x] = csub(sl[x], crs(d, cadd(cmuli(a2), a3)));
                    }
                }
                for (int64_t x = 0; x < nkx; x++) o1[x] = cadd(o1[x], sl[x]);
                for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                for (int k = 0; k < nkr; k++) {
                    const double d = DM(DmS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++) {
                        cplx a1;
                        if (iO > 0) a1 = FB(v, x, k, iO - 1, 0);
                        else if (nkO >= 1) a1 = cneg(cconj(FB(v, refl(x, nkx), k, 1, 0)));
                        else a1 = cmk(0, 0);
                        sl[x] = cadd(sl[x], crs(d, a1));
                    }
                }
                for (int64_t x = 0; x < nkx; x++) {
                    o2[x] = cadd(o2[x], cmuli(sl[x]));
                    o3[x] = cadd(o3[x], sl[x]);
                }
            }
            free(sl);
        }
    }
}

/* Gradient.  fb_math.f90:96-149.  scl_fb(nkx,nkr,0:nkO) -> vec_fb_loc(...,3). */
void ref_fb_grad(double *loc_, const double *scl_, const double *DpS2S, const double *DmS2S, const double *kx,
                 int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc)
{
    cplx *loc = (cplx *)loc_;
    const cplx *s = (const cplx *)scl_;
    const int nmod = nkO + 1;
    memset(loc, 0, sizeof(cplx) * (size_t)nkx * nkr_loc * nmod * 3);
    for (int iO = 0; iO <= nkO; iO++) {
#pragma omp parallel
        {
            cplx *sl = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
#pragma omp for schedule(static)
            for (int kl = 0; kl < nkr_loc; kl++) {
                cplx *o1 = &loc[(int64_t)nkx * (kl + (int64_t)nkr_loc * (iO + (int64_t)nmod * 0))];
                cplx *o2 = &loc[(int64_t)nkx * (kl + (int64_t)nkr_loc * (iO + (int64_t)nmod * 1))];
                cplx *o3 = &loc[(int64_t)nkx * (kl + (int64_t)nkr_loc * (iO + (int64_t)nmod * 2))];
                const cplx *sm = s + (int64_t)nkx * nkr * iO;
                for (int64_t x = 0; x < nkx; x++) o1[x] = cadd(o1[x], cmuli(crs(kx[x], sm[x + (int64_t)nkx * kl])));
                for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                for (int k = 0; k < nkr; k++) {
                    const double d = DM(DmS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++) {
                        cplx a;
                        if (iO > 0) a = s[x + (int64_t)nkx * (k + (int64_t)nkr * (iO - 1))];
                        else if (nkO >= 1) a = cneg(cconj(s[refl(x, nkx) + (int64_t)nkx * (k + (int64_t)nkr * 1)]));
                        else a = cmk(0, 0);
                        sl[x] = cadd(sl[x], crs(d, a));
                    }
                }
                for (int64_t x = 0; x < nkx; x++) {
                    o2[x] = csub(o2[x], sl[x]);
                    o3[x] = cadd(o3[x], cmuli(sl[x]));
                }
                if (iO < nkO) {
                    for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                    for (int k = 0; k < nkr; k++) {
                        const double d = DM(DpS2S, k, kl, iO);
                        for (int64_t x = 0; x < nkx; x++)
                            sl[x] = cadd(sl[x], crs(d, s[x + (int64_t)nkx * (k + (int64_t)nkr * (iO + 1))]));
                    }
                    for (int64_t x = 0; x < nkx; x++) {
                        o2[x] = cadd(o2[x], sl[x]);
                        o3[x] = cadd(o3[x], cmuli(sl[x]));
                    }
                }
            }
            free(sl);
        }
    }
}

/* Divergence core shared by fb_div (fb_math.f90:151-199) and the first half
 * of fb_graddiv (:201-252).  Writes modes 0..nmod_out-1 of out(nkx,nkr_loc,*)
 * from v(nkx,nkr,0:nkO,3); the i*kx term exists for modes <= nkO. */
static void div_core(cplx *out, const cplx *v, const double *DpS2S, const double *DmS2S, const double *kx,
                     int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc, int nmod_out)
{
    const int nmod = nkO + 1;
    for (int iO = 0; iO < nmod_out; iO++) {
#pragma omp parallel
        {
            cplx *sl = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
#pragma omp for schedule(static)
            for (int kl = 0; kl < nkr_loc; kl++) {
                cplx *o = &out[(int64_t)nkx * (kl + (int64_t)nkr_loc * iO)];
                if (iO <= nkO)
                    for (int64_t x = 0; x < nkx; x++) o[x] = cadd(o[x], cmuli(crs(kx[x], FB(v, x, kl, iO, 0))));
                for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                for (int k = 0; k < nkr; k++) {
                    const double d = DM(DmS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++) {
                        cplx a2, a3;
                        if (iO > 0) { a2 = FB(v, x, k, iO - 1, 1); a3 = FB(v, x, k, iO - 1, 2); }
                        else if (nkO >= 1) {
                            a2 = cneg(cconj(FB(v, refl(x, nkx), k, 1, 1)));
                            a3 = cneg(cconj(FB(v, refl(x, nkx), k, 1, 2)));
                        } else { a2 = cmk(0, 0); a3 = cmk(0, 0); }
                        sl[x] = cadd(sl[x], crs(d, csub(cmuli(a3), a2)));
                    }
                }
                if (iO < nkO) {
                    for (int k = 0; k < nkr; k++) {
                        const double d = DM(DpS2S, k, kl, iO);
                        for (int64_t x = 0; x < nkx; x++)
                            sl[x] = cadd(sl[x], crs(d, cadd(cmuli(FB(v, x, k, iO + 1, 2)), FB(v, x, k, iO + 1, 1))));
                    }
                }
                for (int64_t x = 0; x < nkx; x++) o[x] = cadd(o[x], sl[x]);
            }
            free(sl);
        }
    }
}

/* Divergence.  fb_math.f90:151-199.  vec_fb(...,0:nkO,3) -> scl_fb_loc(nkx,nkr_loc,0:nkO). */
void ref_fb_div(double *scl_loc_, const double *vec_, const double *DpS2S, const double *DmS2S, const double *kx,
                int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc)
{
    cplx *out = (cplx *)scl_loc_;
    memset(out, 0, sizeof(cplx) * (size_t)nkx * nkr_loc * (nkO + 1));
    div_core(out, (const cplx *)vec_, DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr_loc, nkO + 1);
}

/* Gradient of divergence, in place.  fb_math.f90:201-293.  The intermediate
 * scalar has nkO+2 modes; its top mode nkO+1 has no i*kx term (:226-229). */
void ref_fb_graddiv(double *vec_, const double *DpS2S, const double *DmS2S, const double *kx,
                    int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc)
{
    cplx *v = (cplx *)vec_;
    const int nmod = nkO + 1;
    cplx *s = (cplx *)calloc((size_t)nkx * nkr_loc * (nkO + 2), sizeof(cplx));
    div_core(s, v, DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr_loc, nkO + 2);
    memset(v, 0, sizeof(cplx) * (size_t)nkx * nkr * nmod * 3);
    /* second half: gradient of s, with D+ always applied to mode iO+1 (:283-288) */
    for (int iO = 0; iO <= nkO; iO++) {
#pragma omp parallel
        {
            cplx *sl = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
#pragma omp for schedule(static)
            for (int kl = 0; kl < nkr_loc; kl++) {
                cplx *o1 = &FB(v, 0, kl, iO, 0), *o2 = &FB(v, 0, kl, iO, 1), *o3 = &FB(v, 0, kl, iO, 2);
                const cplx *sm = s + (int64_t)nkx * nkr_loc * iO;
                for (int64_t x = 0; x < nkx; x++) o1[x] = cadd(o1[x], cmuli(crs(kx[x], sm[x + (int64_t)nkx * kl])));
                for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                for (int k = 0; k < nkr_loc; k++) {
                    const double d = DM(DmS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++) {
                        cplx a;
                        if (iO > 0) a = s[x + (int64_t)nkx * (k + (int64_t)nkr_loc * (iO - 1))];
                        else if (nkO >= 1) a = cneg(cconj(s[refl(x, nkx) + (int64_t)nkx * (k + (int64_t)nkr_loc * 1)]));
                        else a = cmk(0, 0);
                        sl[x] = cadd(sl[x], crs(d, a));
                    }
                }
                for (int64_t x = 0; x < nkx; x++) {
                    o2[x] = csub(o2[x], sl[x]);
                    o3[x] = cadd(o3[x], cmuli(sl[x]));
                }
                for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                for (int k = 0; k < nkr_loc; k++) {
                    const double d = DM(DpS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++)
                        sl[x] = cadd(sl[x], crs(d, s[x + (int64_t)nkx * (k + (int64_t)nkr_loc * (iO + 1))]));
                }
                for (int64_t x = 0; x < nkx; x++) {
                    o2[x] = cadd(o2[x], sl[x]);
                    o3[x] = cadd(o3[x], cmuli(sl[x]));
                }
            }
            free(sl);
        }
    }
    free(s);
}

/* ------------------------------------------------------------------------- */
/* maxwell_solvers.f90 (nkO = number of modes)                               */
/* ------------------------------------------------------------------------- */

/* PSATD with space charge.  maxwell_solvers.f90:18-60.  C1,C2 real (nkx,nkr,nkO,5). */
void ref_maxwell_push_with_spchrg(double *EG_, const double *j_, const double *gn_, const double *gp_,
                                  const double *C1, const double *C2, int32_t nkx, int32_t nkr, int32_t nkO)
{
    cplx *EG = (cplx *)EG_;
    const cplx *J = (const cplx *)j_, *Gn = (const cplx *)gn_, *Gp = (const cplx *)gp_;
    const int64_t n = (int64_t)nkx * nkr * nkO;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n; e++)
        for (int l = 0; l < 3; l++) {
            const cplx E = EG[e + l * n], G = EG[e + (3 + l) * n];
            const cplx j = J[e + l * n], gn = Gn[e + l * n], gp = Gp[e + l * n];
            const cplx Enew = cadd(cadd(cadd(cadd(crs(C1[e], E), crs(C1[e + n], G)), crs(C1[e + 2 * n], j)),
                                        crs(C1[e + 3 * n], gn)), crs(C1[e + 4 * n], gp));
            const cplx Gnew = cadd(cadd(cadd(cadd(crs(C2[e], E), crs(C2[e + n], G)), crs(C2[e + 2 * n], j)),
                                        crs(C2[e + 3 * n], gn)), crs(C2[e + 4 * n], gp));
            EG[e + (3 + l) * n] = Gnew;
            EG[e + l * n] = Enew;
        }
}

/* PSATD without space charge.  maxwell_solvers.f90:62-96.  The reference
 * declares C1,C2 complex and the driver passes real tables (f2py converts);
 * here they are real (nkx,nkr,nkO,3). */
void ref_maxwell_push_wo_spchrg(double *EG_, const double *j_, const double *C1, const double *C2,
                                int32_t nkx, int32_t nkr, int32_t nkO)
{
    cplx *EG = (cplx *)EG_;
    const cplx *J = (const cplx *)j_;
    const int64_t n = (int64_t)nkx * nkr * nkO;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n; e++)
        for (int l = 0; l < 3; l++) {
            const cplx E = EG[e + l * n], G = EG[e + (3 + l) * n], j = J[e + l * n];
            const cplx Enew = cadd(cadd(crs(C1[e], E), crs(C1[e + n], G)), crs(C1[e + 2 * n], j));
            const cplx Gnew = cadd(cadd(crs(C2[e], E), crs(C2[e + n], G)), crs(C2[e + 2 * n], j));
            EG[e + (3 + l) * n] = Gnew;
            EG[e + l * n] = Enew;
        }
}

/* Static initial push.  maxwell_solvers.f90:98-129.  C1,C2 complex (nkx,nkr,nkO,2). */
void ref_maxwell_init_push(double *EG_, const double *j_, const double *gn_, const double *C1_, const double *C2_,
                           int32_t nkx, int32_t nkr, int32_t nkO)
{
    cplx *EG = (cplx *)EG_;
    const cplx *J = (const cplx *)j_, *Gn = (const cplx *)gn_, *C1 = (const cplx *)C1_, *C2 = (const cplx *)C2_;
    const int64_t n = (int64_t)nkx * nkr * nkO;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n; e++)
        for (int l = 0; l < 3; l++) {
            EG[e + l * n] = cadd(cadd(EG[e + l * n], cmul(C1[e], J[e + l * n])), cmul(C1[e + n], Gn[e + l * n]));
            EG[e + (3 + l) * n] =
                cadd(cadd(EG[e + (3 + l) * n], cmul(C2[e], J[e + l * n])), cmul(C2[e + n], Gn[e + l * n]));
        }
}

/* Poisson correction of the current.  maxwell_solvers.f90:131-164. */
void ref_poiss_corr(double *j_, const double *gdj_, const double *gn_, const double *gp_, double dt_inv,
                    const double *w2_inv, int32_t nkx, int32_t nkr, int32_t nkO)
{
    cplx *J = (cplx *)j_;
    const cplx *GD = (const cplx *)gdj_, *Gn = (const cplx *)gn_, *Gp = (const cplx *)gp_;
    const int64_t n = (int64_t)nkx * nkr * nkO;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n; e++)
        for (int l = 0; l < 3; l++) {
            const cplx d = crs(dt_inv, csub(Gp[e + l * n], Gn[e + l * n]));
            J[e + l * n] = cadd(J[e + l * n], crs(w2_inv[e], cadd(GD[e + l * n], d)));
        }
}

/* Poisson correction of the static kick.  maxwell_solvers.f90:166-197.  DT complex (nkx). */
void ref_poiss_corr_stat(double *j_, const double *gdj_, const double *gn_, const double *DT_, const double *w2_inv,
                         int32_t nkx, int32_t nkr, int32_t nkO)
{
    cplx *J = (cplx *)j_;
    const cplx *GD = (const cplx *)gdj_, *Gn = (const cplx *)gn_, *DT = (const cplx *)DT_;
    const int64_t n = (int64_t)nkx * nkr * nkO;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n; e++)
        for (int l = 0; l < 3; l++)
            J[e + l * n] = cadd(J[e + l * n], crs(w2_inv[e], cadd(GD[e + l * n], cmul(DT[e % nkx], Gn[e + l * n]))));
}

/* Free drift over one step.  maxwell_solvers.f90:199-226: PROPGTR = exp(fact*kx) with
 * fact = -0.5*i*dt*beta0, whose real part is zero: (cos, sin) of ((-0.5*dt)*beta0)*kx. */
void ref_field_drift(double *EG_, const double *kx, double beta0, double dt, int32_t nkx, int32_t nkr, int32_t nkO)
{
    cplx *EG = (cplx *)EG_;
    const int64_t n = (int64_t)nkx * nkr * nkO;
    const double a = (-0.5 * dt) * beta0;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n; e++) {
        const double ph = a * kx[e % nkx];
        const cplx p = cmk(cos(ph), sin(ph));
        for (int l = 0; l < 6; l++) EG[e + l * n] = cmul(EG[e + l * n], p);
    }
}

/* Elementwise helpers.  maxwell_solvers.f90:228-318.  ncomp = 3 (vec) or 1 (scl). */
void ref_omp_mult(double *v_, const double *A, int32_t nkx, int32_t nkr, int32_t nkO, int32_t ncomp)
{
    cplx *v = (cplx *)v_;
    const int64_t n = (int64_t)nkx * nkr * nkO;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n; e++)
        for (int l = 0; l < ncomp; l++) v[e + l * n] = crs(A[e], v[e + l * n]);
}

void ref_omp_add(double *v_, const double *A_, int32_t nkx, int32_t nkr, int32_t nkO, int32_t ncomp)
{
    cplx *v = (cplx *)v_;
    const cplx *A = (const cplx *)A_;
    const int64_t n = (int64_t)nkx * nkr * nkO;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n * ncomp; e++) v[e] = cadd(v[e], A[e]);
}

/* ------------------------------------------------------------------------- */
/* devices.f90                                                               */
/* ------------------------------------------------------------------------- */

/* Analytic planar undulator.  devices.f90:162-203.  params = (a0, lambda, X0, Lx). */
/* Linearly tapered analytic undulator.  devices.f90:117-160: the amplitude of
 * Undul_Analytic times 1 + taper (x - X0 - 0.5 Lx) / (0.5 Lx). */
void ref_undul_analytic_taper(const double *coord, double *Fld, double t, const double *params, int32_t np)
{
    (void)t;
    const double pi = 4.0 * atan(1.0);
    const double a0 = params[0], lambda = params[1], X0 = params[2], Lx = params[3], taper = params[4];
    const double ku = 2.0 * pi / lambda;
#pragma omp parallel for schedule(static)
    for (int64_t ip = 0; ip < np; ip++) {
        const double x = coord[3 * ip], y = coord[3 * ip + 1];
        double ampl;
        if (x <= X0 || x >= X0 + Lx) ampl = 0.0;
        else if (x > X0 && x < X0 + lambda) ampl = (x - X0) / lambda;
        else if (x > X0 + Lx - lambda && x < X0 + Lx) ampl = (X0 + Lx - x) / lambda;
        else ampl = 1.0;
        ampl = ampl * (1.0 + taper * (x - X0 - 0.5 * Lx) / (0.5 * Lx));
        ampl = ampl * a0;
        Fld[6 * ip + 4] = Fld[6 * ip + 4] + ampl * sin(ku * (x - X0)) * cosh(ku * y);
        Fld[6 * ip + 3] = Fld[6 * ip + 3] + ampl * cos(ku * (x - X0)) * sinh(ku * y);
    }
}

/* Mapped undulator.  devices.f90:18-62 (tap = 0, params = lambda, Xleft, dx) and
 * 64-115 (tap = 1, params add Lx, taper).  a0(2,nx) Fortran order, nodes 1..nx;
 * a node outside 1..nx contributes zero (the reference reads before the map there). */
static void undul_mapped(const double *coord, double *Fld, const double *a0, const double *params, int32_t np,
                         int32_t nx, int tap)
{
    const double pi = 4.0 * atan(1.0);
    const double lambda = params[0], Xleft = params[1], dx = params[2];
    const double Lx = tap ? params[3] : 0.0, taper = tap ? params[4] : 0.0;
    const double ku = 2.0 * pi / lambda, dx_inv = 1.0 / dx, Xright = Xleft + nx * dx;
    const double x_shift = 0.5 * (nx * dx - Lx);
#pragma omp parallel for schedule(static)
    for (int64_t ip = 0; ip < np; ip++) {
        const double xp = coord[3 * ip];
        if (xp < Xleft + dx || xp > Xright - dx) continue;
        const double yp = coord[3 * ip + 1];
        const double fx = floor((xp - Xleft) * dx_inv + 0.5);
        if (!(fx >= 0.0 && fx <= (double)nx + 1.0)) continue;
        const int ix = (int)fx;
        const double ddx = (xp - Xleft) * dx_inv - ix;
        const double S0[3] = {0.5 * ((0.5 - ddx) * (0.5 - ddx)), 0.75 - ddx * ddx, 0.5 * ((0.5 + ddx) * (0.5 + ddx))};
        double b1 = 0.0, b2 = 0.0;
        for (int k = 0; k < 3; k++) {
            const int j = ix - 1 + k;
            const int in = j >= 1 && j <= nx;
            b1 = b1 + S0[k] * (in ? a0[2 * (int64_t)(j - 1)] : 0.0);
            b2 = b2 + S0[k] * (in ? a0[2 * (int64_t)(j - 1) + 1] : 0.0);
        }
        if (tap) {
            const double amp = 1.0 + taper * (xp - x_shift - Xleft - 0.5 * Lx) / (0.5 * Lx);
            Fld[6 * ip + 4] = Fld[6 * ip + 4] + amp * b1 * cosh(ku * yp);
            Fld[6 * ip + 3] = Fld[6 * ip + 3] + amp * b2 * sinh(ku * yp);
        } else {
            Fld[6 * ip + 4] = Fld[6 * ip + 4] + b1 * cosh(ku * yp);
            Fld[6 * ip + 3] = Fld[6 * ip + 3] + b2 * sinh(ku * yp);
        }
    }
}

void ref_undul_mapped(const double *coord, double *Fld, double t, const double *a0, const double *params, int32_t np,
                      int32_t nx)
{
    (void)t;
    undul_mapped(coord, Fld, a0, params, np, nx, 0);
}

void ref_undul_mapped_tap(const double *coord, double *Fld, double t, const double *a0, const double *params,
                          int32_t np, int32_t nx)
{
    (void)t;
    undul_mapped(coord, Fld, a0, params, np, nx, 1);
}

/* Plane wave.  devices.f90:205-252. */
void ref_planewave(const double *coord, double *Fld, double t, const double *params, int32_t np)
{
    const double pi = 4.0 * atan(1.0);
    const double a0 = params[0], lambda = params[1], X0 = params[2], Lx = params[3], ramp = params[4];
    const double theta = params[5], phi0 = params[6];
    const double sinth = sin(theta), costh = cos(theta), k0 = 2.0 * pi / lambda;
#pragma omp parallel for schedule(static)
    for (int64_t ip = 0; ip < np; ip++) {
        const double x = coord[3 * ip], y = coord[3 * ip + 1];
        double ampl;
        if (x <= X0 || x >= X0 + Lx) ampl = 0.0;
        else if (x > X0 && x < X0 + ramp) ampl = (x - X0) / ramp;
        else if (x > X0 + Lx - ramp && x < X0 + Lx) ampl = (X0 + Lx - x) / ramp;
        else ampl = 1.0;
        ampl = a0 * ampl * sin(k0 * (x * costh + y * sinth - t) + phi0);
        Fld[6 * ip + 2] = Fld[6 * ip + 2] + ampl;
        Fld[6 * ip + 3] = Fld[6 * ip + 3] + ampl * sinth;
        Fld[6 * ip + 4] = Fld[6 * ip + 4] - ampl * costh;
    }
}

/* Gaussian beam.  devices.f90:254-297. */
void ref_gaussbeam(const double *coord, double *Fld, double time, double a0, const double *params, int32_t np)
{
    const double pi = 4.0 * atan(1.0);
    const double lambda = params[0], axis = params[1], x0 = params[2], y0 = params[3], z0 = params[4];
    const double Lx2_inv = 1.0 / (params[5] * params[5]), Ly2_inv = 1.0 / (params[6] * params[6]);
    const double Lz2_inv = 1.0 / (params[7] * params[7]);
#pragma omp parallel for schedule(static)
    for (int64_t ip = 0; ip < np; ip++) {
        const double xp = coord[3 * ip] - x0 - axis * time, yp = coord[3 * ip + 1] - y0, zp = coord[3 * ip + 2] - z0;
        const double e = a0 * exp(-(xp * xp * Lx2_inv) - yp * yp * Ly2_inv - zp * zp * Lz2_inv) * sin(2.0 * pi / lambda * xp);
        Fld[6 * ip + 2] = Fld[6 * ip + 2] + e;
        Fld[6 * ip + 4] = Fld[6 * ip + 4] - axis * e;
    }
}

void ref_undul_analytic(const double *coord, double *Fld, double t, const double *params, int32_t np)
{
    (void)t;
    const double pi = 4.0 * atan(1.0);
    const double a0 = params[0], lambda = params[1], X0 = params[2], Lx = params[3];
    const double ku = 2.0 * pi / lambda;
#pragma omp parallel for schedule(static)
    for (int64_t ip = 0; ip < np; ip++) {
        const double x = coord[3 * ip], y = coord[3 * ip + 1];
        double ampl;
        if (x <= X0 || x >= X0 + Lx) ampl = 0.0;
        else if (x > X0 && x < X0 + lambda) ampl = (x - X0) / lambda;
        else if (x > X0 + Lx - lambda && x < X0 + Lx) ampl = (X0 + Lx - x) / lambda;
        else ampl = 1.0;
        ampl = ampl * a0;
        Fld[6 * ip + 4] = Fld[6 * ip + 4] + ampl * sin(ku * (x - X0)) * cosh(ku * y);
        Fld[6 * ip + 3] = Fld[6 * ip + 3] + ampl * cos(ku * (x - X0)) * sinh(ku * y);
    }
}

/* ------------------------------------------------------------------------- */
/* grid_deps_env.f90: deposits and gather of the envelope (KxShift) solver   */
/* ------------------------------------------------------------------------- */
/* Mode arrays run over -nkO..nkO: mode iO sits at index iO + nkO.  The
 * particle's weight carries the carrier phase e^{-i kx0 x} (deposits) or its
 * conjugate (gather). */

/* Cell, weights and phase powers of an envelope particle: the deposits use
 * (y - i z)/r and 0 on the axis, the gather (y + i z)/r and 1 on the axis
 * (grid_deps_env.f90:56-60, :131-135, :203-207).  Particles whose cell lies
 * outside the arrays are skipped, as for ref_geom. */
static inline int ref_geom_env(pgeom *g, double xp, double yp, double zp, double rp, double leftX,
                               const double *Rgrid, double dx_inv, double dr_inv, int nx, int nr, int nkO, int gather)
{
    const double fx = (xp - leftX) * dx_inv, fr = (rp - Rgrid[0]) * dr_inv;
    if (!(fx >= 0.0 && fx < (double)nx) || !(fr >= 0.0 && fr < (double)nr)) return 0;
    g->ix = (int64_t)floor(fx);
    g->ir = (int64_t)floor(fr);
    g->S0[1][0] = (xp - leftX) * dx_inv - (double)g->ix;
    g->S0[0][0] = 1.0 - g->S0[1][0];
    g->S0[1][1] = (rp - Rgrid[g->ir]) * dr_inv;
    g->S0[0][1] = 1.0 - g->S0[1][1];
    cplx ph;
    if (rp > 0.0) ph = gather ? cmk(yp / rp, zp / rp) : cmk(yp / rp, -zp / rp);
    else ph = gather ? cmk(1.0, 0.0) : cmk(0.0, 0.0);
    g->phase[0] = cmk(1.0, 0.0);
    for (int iO = 1; iO <= nkO; iO++) g->phase[iO] = cmul(g->phase[iO - 1], ph);
    return 1;
}

/* Envelope current.  grid_deps_env.f90:18-95: only the third (x... the
 * reference's l = 3) component is deposited (`do l=3,3`), with the velocity
 * u/gamma and the complex weight w e^{-i kx0 x}; mode -iO gets the conjugate
 * phase; then the ghost fold of all three components.
 * curr(0:nx,0:nr,-nkO:nkO,3). */
void ref_dep_curr_env(const double *coord, const double *momenta, const double *wghts, double *curr_,
                      double leftX, const double *Rgrid, double dx_inv, double dr_inv, double kx0,
                      int32_t np, int32_t nx, int32_t nr, int32_t nkO)
{
    cplx *curr = (cplx *)curr_;
    const int nm = 2 * nkO + 1;
    const int64_t sr = nx + 1, sm = (int64_t)(nx + 1) * (nr + 1), sl = sm * nm;
    for (int64_t ip = 0; ip < np; ip++) {
        const double w = wghts[ip];
        if (w == 0.0) continue;
        const double xp = coord[3 * ip], yp = coord[3 * ip + 1], zp = coord[3 * ip + 2];
        double v[3] = {momenta[3 * ip], momenta[3 * ip + 1], momenta[3 * ip + 2]};
        const double rp = sqrt(yp * yp + zp * zp);
        if (rp >= Rgrid[nr] || (fabs(v[0]) + fabs(v[1]) + fabs(v[2])) == 0.0) continue;
        const double gp = sqrt(1.0 + v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int c = 0; c < 3; c++) v[c] = v[c] / gp;
        const cplx wp = cmul(cmk(w, 0.0), cmk(cos(xp * kx0), -sin(xp * kx0)));
        pgeom g;
        if (!ref_geom_env(&g, xp, yp, zp, rp, leftX, Rgrid, dx_inv, dr_inv, nx, nr, nkO, 0)) continue;
        cplx cp[2][2];
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) cp[a][b] = cmul(cmk(g.S0[a][0] * g.S0[b][1], 0.0), wp);
        const int l = 2;
        for (int iO = 0; iO <= nkO; iO++) {
            const cplx pv = cmul(g.phase[iO], cmk(v[l], 0.0));
            const cplx pc = cmul(cconj(g.phase[iO]), cmk(v[l], 0.0));
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) {
                    cplx *dst = curr + (g.ix + a) + (g.ir + b) * sr + (iO + nkO) * sm + l * sl;
                    *dst = cadd(*dst, cmul(pv, cp[a][b]));
                }
            if (iO > 0)
                for (int b = 0; b < 2; b++)
                    for (int a = 0; a < 2; a++) {
                        cplx *dst = curr + (g.ix + a) + (g.ir + b) * sr + (nkO - iO) * sm + l * sl;
                        *dst = cadd(*dst, cmul(pc, cp[a][b]));
                    }
        }
    }
    for (int l = 0; l < 3; l++)
        for (int m = 0; m < nm; m++)
            for (int64_t ix = 0; ix <= nx; ix++) {
                cplx *r0 = curr + ix + m * sm + l * sl;
                r0[sr] = csub(r0[sr], r0[0]);
                r0[0] = cmk(0.0, 0.0);
            }
}

/* Envelope charge.  grid_deps_env.f90:97-162, the complex weight applied
 * twice (grid_deps_env.f90:145-147).  dens(0:nx,0:nr,-nkO:nkO). */
void ref_dep_dens_env(const double *coord, const double *wghts, double *dens_, double leftX, const double *Rgrid,
                      double dx_inv, double dr_inv, double kx0, int32_t np, int32_t nx, int32_t nr, int32_t nkO)
{
    cplx *dens = (cplx *)dens_;
    const int nm = 2 * nkO + 1;
    const int64_t sr = nx + 1, sm = (int64_t)(nx + 1) * (nr + 1);
    for (int64_t ip = 0; ip < np; ip++) {
        const double w = wghts[ip];
        if (w == 0.0) continue;
        const double xp = coord[3 * ip], yp = coord[3 * ip + 1], zp = coord[3 * ip + 2];
        const double rp = sqrt(yp * yp + zp * zp);
        if (rp >= Rgrid[nr]) continue;
        const cplx wp = cmul(cmk(w, 0.0), cmk(cos(xp * kx0), -sin(xp * kx0)));
        pgeom g;
        if (!ref_geom_env(&g, xp, yp, zp, rp, leftX, Rgrid, dx_inv, dr_inv, nx, nr, nkO, 0)) continue;
        cplx dp[2][2];
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) {
                dp[a][b] = cmul(cmk(g.S0[a][0] * g.S0[b][1], 0.0), wp);
                dp[a][b] = cmul(dp[a][b], wp);
            }
        for (int iO = 0; iO <= nkO; iO++) {
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) {
                    cplx *dst = dens + (g.ix + a) + (g.ir + b) * sr + (iO + nkO) * sm;
                    *dst = cadd(*dst, cmul(g.phase[iO], dp[a][b]));
                }
            if (iO > 0)
                for (int b = 0; b < 2; b++)
                    for (int a = 0; a < 2; a++) {
                        cplx *dst = dens + (g.ix + a) + (g.ir + b) * sr + (nkO - iO) * sm;
                        *dst = cadd(*dst, cmul(cconj(g.phase[iO]), dp[a][b]));
                    }
        }
    }
    for (int m = 0; m < nm; m++)
        for (int64_t ix = 0; ix <= nx; ix++) {
            cplx *r0 = dens + ix + m * sm;
            r0[sr] = csub(r0[sr], r0[0]);
            r0[0] = cmk(0.0, 0.0);
        }
}

/* Envelope gather.  grid_deps_env.f90:164-238: the interpolation weights
 * times e^{+i kx0 x} and the phase powers (conjugated for mode -iO); the real
 * part of each product, summed per mode (mode +iO, then -iO) and added to
 * Fld_tot.  Fld(0:nx,0:nr,-nkO:nkO,6). */
void ref_proj_fld_env(const double *coord, const double *wghts, const double *Fld_, double *Fld_tot,
                      double leftX, const double *Rgrid, double dx_inv, double dr_inv, double kx0,
                      int32_t np, int32_t nx, int32_t nr, int32_t nkO)
{
    const cplx *Fld = (const cplx *)Fld_;
    const int nm = 2 * nkO + 1;
    const int64_t sr = nx + 1, sm = (int64_t)(nx + 1) * (nr + 1), sl = sm * nm;
#pragma omp parallel for schedule(static)
    for (int64_t ip = 0; ip < np; ip++) {
        const double wp = wghts[ip];
        if (wp == 0.0) continue;
        const double xp = coord[3 * ip], yp = coord[3 * ip + 1], zp = coord[3 * ip + 2];
        const double rp = sqrt(yp * yp + zp * zp);
        if (rp >= Rgrid[nr]) continue;
        pgeom g;
        if (!ref_geom_env(&g, xp, yp, zp, rp, leftX, Rgrid, dx_inv, dr_inv, nx, nr, nkO, 1)) continue;
        const cplx carrier = cmk(cos(xp * kx0), sin(xp * kx0));
        cplx pc[2][2];
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) pc[a][b] = cmul(cmk(g.S0[b][1] * g.S0[a][0], 0.0), carrier);
        double Fp[6] = {0, 0, 0, 0, 0, 0};
        for (int iO = 0; iO <= nkO; iO++) {
            for (int sgn = 0; sgn < (iO > 0 ? 2 : 1); sgn++) {
                const cplx ph = sgn ? cconj(g.phase[iO]) : g.phase[iO];
                const int m = sgn ? nkO - iO : nkO + iO;
                cplx po[2][2];
                for (int b = 0; b < 2; b++)
                    for (int a = 0; a < 2; a++) po[a][b] = cmul(pc[a][b], ph);
                for (int l = 0; l < 6; l++) {
                    double s = 0.0;
                    for (int b = 0; b < 2; b++)
                        for (int a = 0; a < 2; a++) {
                            const cplx f = Fld[(g.ix + a) + (g.ir + b) * sr + m * sm + l * sl];
                            s += po[a][b].re * f.re - po[a][b].im * f.im;
                        }
                    Fp[l] = Fp[l] + s;
                }
            }
        }
        for (int l = 0; l < 6; l++) Fld_tot[6 * ip + l] = Fld_tot[6 * ip + l] + Fp[l];
    }
}

/* grid_deps_env.f90:240-283: every mode times 1/pi; the ghost row is +row 1
 * when nkO == 0 and -row 1 for every mode otherwise.  eb(0:nx,0:nr,-nkO:nkO,6). */
void ref_eb_correction_env(double *eb_, int32_t nx, int32_t nr, int32_t nkO)
{
    cplx *eb = (cplx *)eb_;
    const double pi = 4.0 * atan(1.0);
    const double pi_inv = 1. / pi;
    const int nm = 2 * nkO + 1;
    const int64_t sr = nx + 1, sm = (int64_t)(nx + 1) * (nr + 1), n = sm * nm * 6;
    for (int64_t j = 0; j < n; j++) eb[j] = crs(pi_inv, eb[j]);
    for (int64_t c = 0; c < (int64_t)nm * 6; c++)
        for (int64_t ix = 0; ix <= nx; ix++) {
            cplx *col = eb + c * sm + ix;
            col[0] = nkO == 0 ? col[sr] : cneg(col[sr]);
        }
}

/* ------------------------------------------------------------------------- */
/* fb_math_env.f90: FB-space operators of the envelope solver                */
/* ------------------------------------------------------------------------- */
/* Fields (nkx, nkr, -nkO:nkO, c); D+-(nkr, nkr_loc, -nkO-1:nkO+1).  Mode m at
 * field index m + nkO, at D index m + nkO + 1.  The coupling matrices apply
 * only between modes present on the grid (no reflected mode). */
#define FE(a, x, k, m, c) (a)[(x) + (int64_t)nkx * ((k) + (int64_t)nkr * ((m) + nkO + (int64_t)nm * (c)))]
#define FL(a, x, k, m, c) (a)[(x) + (int64_t)nkx * ((k) + (int64_t)nkr_loc * ((m) + nkO + (int64_t)nm * (c)))]
#define DE(D, k, kl, m) (D)[(k) + (int64_t)nkr * ((kl) + (int64_t)nkr_loc * ((m) + nkO + 1))]

/* Gradient.  fb_math_env.f90:18-61. */
void ref_fb_grad_env(double *loc_, const double *scl_, const double *DpS2S, const double *DmS2S, const double *kx,
                     int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc)
{
    cplx *loc = (cplx *)loc_;
    const cplx *s = (const cplx *)scl_;
    const int nm = 2 * nkO + 1;
    memset(loc, 0, sizeof(cplx) * (size_t)nkx * nkr_loc * nm * 3);
    cplx *sl = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
    for (int iO = -nkO; iO <= nkO; iO++)
        for (int kl = 0; kl < nkr_loc; kl++) {
            for (int64_t x = 0; x < nkx; x++)
                FL(loc, x, kl, iO, 0) = cadd(FL(loc, x, kl, iO, 0), cmuli(crs(kx[x], FE(s, x, kl, iO, 0))));
            if (iO > -nkO) {
                for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                for (int k = 0; k < nkr; k++) {
                    const double d = DE(DmS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++) sl[x] = cadd(sl[x], crs(d, FE(s, x, k, iO - 1, 0)));
                }
                for (int64_t x = 0; x < nkx; x++) {
                    FL(loc, x, kl, iO, 1) = csub(FL(loc, x, kl, iO, 1), sl[x]);
                    FL(loc, x, kl, iO, 2) = cadd(FL(loc, x, kl, iO, 2), cmuli(sl[x]));
                }
            }
            if (iO < nkO) {
                for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                for (int k = 0; k < nkr; k++) {
                    const double d = DE(DpS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++) sl[x] = cadd(sl[x], crs(d, FE(s, x, k, iO + 1, 0)));
                }
                for (int64_t x = 0; x < nkx; x++) {
                    FL(loc, x, kl, iO, 1) = cadd(FL(loc, x, kl, iO, 1), sl[x]);
                    FL(loc, x, kl, iO, 2) = cadd(FL(loc, x, kl, iO, 2), cmuli(sl[x]));
                }
            }
        }
    free(sl);
}

/* Divergence core of fb_div_env (fb_math_env.f90:63-104) for modes
 * mlo..mhi of the output (nkr_loc rows, mode m at index m - mlo of nmo), the
 * input modes -nkO..nkO; the i kx term only where the output mode exists in
 * the input.  Shared with the first half of fb_graddiv_env. */
static void div_env_core(cplx *out, int mlo, int mhi, const cplx *v, const double *DpS2S, const double *DmS2S,
                         const double *kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc)
{
    const int nm = 2 * nkO + 1, nmo = mhi - mlo + 1;
    cplx *sl = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
#define FO(x, kl, m) out[(x) + (int64_t)nkx * ((kl) + (int64_t)nkr_loc * ((m) - mlo))]
    for (int iO = mlo; iO <= mhi; iO++)
        for (int kl = 0; kl < nkr_loc; kl++) {
            if (iO >= -nkO && iO <= nkO)
                for (int64_t x = 0; x < nkx; x++)
                    FO(x, kl, iO) = cadd(FO(x, kl, iO), cmuli(crs(kx[x], FE(v, x, kl, iO, 0))));
            for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
            if (iO > -nkO)
                for (int k = 0; k < nkr; k++) {
                    const double d = DE(DmS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++)
                        sl[x] = csub(sl[x], crs(d, csub(FE(v, x, k, iO - 1, 1), cmuli(FE(v, x, k, iO - 1, 2)))));
                }
            if (iO < nkO)
                for (int k = 0; k < nkr; k++) {
                    const double d = DE(DpS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++)
                        sl[x] = cadd(sl[x], crs(d, cadd(FE(v, x, k, iO + 1, 1), cmuli(FE(v, x, k, iO + 1, 2)))));
                }
            for (int64_t x = 0; x < nkx; x++) FO(x, kl, iO) = cadd(FO(x, kl, iO), sl[x]);
        }
#undef FO
    free(sl);
    (void)nmo;
    (void)nm;
}

/* Divergence.  fb_math_env.f90:63-104. */
void ref_fb_div_env(double *scl_loc_, const double *vec_, const double *DpS2S, const double *DmS2S, const double *kx,
                    int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc)
{
    cplx *out = (cplx *)scl_loc_;
    memset(out, 0, sizeof(cplx) * (size_t)nkx * nkr_loc * (2 * nkO + 1));
    div_env_core(out, -nkO, nkO, (const cplx *)vec_, DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr_loc);
}

/* Curl.  fb_math_env.f90:106-162.  In the m-1 branch the reference forms the
 * x-component term -D-.(i v_y + v_z) and then discards it (sliceKR is reset
 * before it is used, fb_math_env.f90:141-146); here it is not formed. */
void ref_fb_rot_env(double *loc_, const double *vec_, const double *DpS2S, const double *DmS2S, const double *kx,
                    int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc)
{
    cplx *loc = (cplx *)loc_;
    const cplx *v = (const cplx *)vec_;
    const int nm = 2 * nkO + 1;
    memset(loc, 0, sizeof(cplx) * (size_t)nkx * nkr_loc * nm * 3);
    cplx *sl = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
    for (int iO = -nkO; iO <= nkO; iO++)
        for (int kl = 0; kl < nkr_loc; kl++) {
            for (int64_t x = 0; x < nkx; x++) {
                FL(loc, x, kl, iO, 1) = csub(FL(loc, x, kl, iO, 1), cmuli(crs(kx[x], FE(v, x, kl, iO, 2))));
                FL(loc, x, kl, iO, 2) = cadd(FL(loc, x, kl, iO, 2), cmuli(crs(kx[x], FE(v, x, kl, iO, 1))));
            }
            if (iO < nkO) {
                for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                for (int k = 0; k < nkr; k++) {
                    const double d = DE(DpS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++)
                        sl[x] = csub(sl[x], crs(d, csub(cmuli(FE(v, x, k, iO + 1, 1)), FE(v, x, k, iO + 1, 2))));
                }
                for (int64_t x = 0; x < nkx; x++) FL(loc, x, kl, iO, 0) = cadd(FL(loc, x, kl, iO, 0), sl[x]);
                for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                for (int k = 0; k < nkr; k++) {
                    const double d = DE(DpS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++) sl[x] = cadd(sl[x], crs(d, FE(v, x, k, iO + 1, 0)));
                }
                for (int64_t x = 0; x < nkx; x++) {
                    FL(loc, x, kl, iO, 1) = cadd(FL(loc, x, kl, iO, 1), cmuli(sl[x]));
                    FL(loc, x, kl, iO, 2) = csub(FL(loc, x, kl, iO, 2), sl[x]);
                }
            }
            if (iO > -nkO) {
                for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
                for (int k = 0; k < nkr; k++) {
                    const double d = DE(DmS2S, k, kl, iO);
                    for (int64_t x = 0; x < nkx; x++) sl[x] = cadd(sl[x], crs(d, FE(v, x, k, iO - 1, 0)));
                }
                for (int64_t x = 0; x < nkx; x++) {
                    FL(loc, x, kl, iO, 1) = cadd(FL(loc, x, kl, iO, 1), cmuli(sl[x]));
                    FL(loc, x, kl, iO, 2) = cadd(FL(loc, x, kl, iO, 2), sl[x]);
                }
            }
        }
    free(sl);
}

/* Gradient of the divergence, in place.  fb_math_env.f90:164-233: the
 * divergence on modes -nkO-1..nkO+1 (the outer two from the coupling terms
 * alone), then its gradient on modes -nkO..nkO, each mode taking D- of mode
 * m-1 and D+ of mode m+1 unconditionally. */
void ref_fb_graddiv_env(double *vec_, const double *DpS2S, const double *DmS2S, const double *kx,
                        int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc)
{
    cplx *v = (cplx *)vec_;
    const int nm = 2 * nkO + 1, nmx = nm + 2;
    cplx *s = (cplx *)calloc((size_t)nkx * nkr_loc * nmx, sizeof(cplx));
    div_env_core(s, -nkO - 1, nkO + 1, v, DpS2S, DmS2S, kx, nkx, nkr, nkO, nkr_loc);
#define FS(x, k, m) s[(x) + (int64_t)nkx * ((k) + (int64_t)nkr_loc * ((m) + nkO + 1))]
    memset(v, 0, sizeof(cplx) * (size_t)nkx * nkr * nm * 3);
    cplx *sl = (cplx *)malloc(sizeof(cplx) * (size_t)nkx);
    for (int iO = -nkO; iO <= nkO; iO++)
        for (int kl = 0; kl < nkr_loc; kl++) {
            for (int64_t x = 0; x < nkx; x++) FE(v, x, kl, iO, 0) = cadd(FE(v, x, kl, iO, 0), cmuli(crs(kx[x], FS(x, kl, iO))));
            for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
            for (int k = 0; k < nkr; k++) {
                const double d = DE(DmS2S, k, kl, iO);
                for (int64_t x = 0; x < nkx; x++) sl[x] = cadd(sl[x], crs(d, FS(x, k, iO - 1)));
            }
            for (int64_t x = 0; x < nkx; x++) {
                FE(v, x, kl, iO, 1) = csub(FE(v, x, kl, iO, 1), sl[x]);
                FE(v, x, kl, iO, 2) = cadd(FE(v, x, kl, iO, 2), cmuli(sl[x]));
            }
            for (int64_t x = 0; x < nkx; x++) sl[x] = cmk(0, 0);
            for (int k = 0; k < nkr; k++) {
                const double d = DE(DpS2S, k, kl, iO);
                for (int64_t x = 0; x < nkx; x++) sl[x] = cadd(sl[x], crs(d, FS(x, k, iO + 1)));
            }
            for (int64_t x = 0; x < nkx; x++) {
                FE(v, x, kl, iO, 1) = cadd(FE(v, x, kl, iO, 1), sl[x]);
                FE(v, x, kl, iO, 2) = cadd(FE(v, x, kl, iO, 2), cmuli(sl[x]));
            }
        }
#undef FS
    free(sl);
    free(s);
}
#undef FE
#undef FL
#undef DE

/* maxwell_push_wo_spchrg with complex coefficient tables, as the envelope
 * solver passes them (maxwell_solvers.f90:62-96; solvers.py:244-267).
 * C1, C2 complex (nkx,nkr,nkO,3); nkO = the number of modes. */
void ref_maxwell_push_wo_spchrg_cplx(double *EG_, const double *j_, const double *C1_, const double *C2_,
                                     int32_t nkx, int32_t nkr, int32_t nkO)
{
    cplx *EG = (cplx *)EG_;
    const cplx *J = (const cplx *)j_, *C1 = (const cplx *)C1_, *C2 = (const cplx *)C2_;
    const int64_t n = (int64_t)nkx * nkr * nkO;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n; e++)
        for (int l = 0; l < 3; l++) {
            const cplx E = EG[e + l * n], G = EG[e + (3 + l) * n], j = J[e + l * n];
            const cplx Enew = cadd(cadd(cmul(C1[e], E), cmul(C1[e + n], G)), cmul(C1[e + 2 * n], j));
            const cplx Gnew = cadd(cadd(cmul(C2[e], E), cmul(C2[e + n], G)), cmul(C2[e + 2 * n], j));
            EG[e + (3 + l) * n] = Gnew;
            EG[e + l * n] = Enew;
        }
}
