/*
 * chimera_hip.h -- C ABI of libchimera_hip, the MI355X (gfx950) implementation
 * of CHIMERA's particle-in-cell step.
 *
 * Drop-in tier: one entry point per routine of the reference's f2py module
 * `fimera` on the PIC-step path.  The argument list of each is the Fortran
 * subroutine's dummy-argument list, in the same order and with the same
 * meaning (including the dimension arguments f2py hides, with the meaning the
 * Fortran declaration gives them: e.g. `nx` is the last index of a 0:nx array,
 * and `nkO` is the MAXIMUM mode in grid_deps/fb_math routines but the NUMBER
 * of modes in fb_io/maxwell_solvers routines).  Scalars are passed by value.
 * Arrays are Fortran-ordered; complex(kind=8) arrays are interleaved (re, im)
 * doubles passed as `double*`.
 *
 * Array pointers may point to host or to device memory, argument by argument:
 * device pointers are used in place (the call is then asynchronous on the
 * library stream), host pointers are staged through device buffers and the
 * call returns when their results are back in host memory.  Scalar outputs
 * (num2stay, GoOut, indPart, IndInChnk) are host pointers.
 *
 * Every entry point returns 0 on success or a negative status
 * (CHIMERA_EINVAL, CHIMERA_EHIP, CHIMERA_ENOMEM, CHIMERA_EUNSUP) and never
 * aborts; chimera_last_error() describes the last failure of the calling thread.
 *
 * Reference file:line citations are relative to the reference repository
 * (hightower8083/chimera).
 */
#ifndef CHIMERA_HIP_H
#define CHIMERA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHIMERA_OK 0
#define CHIMERA_EINVAL (-1)
#define CHIMERA_EHIP (-2)
#define CHIMERA_ENOMEM (-3)
#define CHIMERA_EUNSUP (-4)

/* ---------------------------------------------------------------- runtime */
const char* chimera_version(void);
const char* chimera_last_error(void);
/* CHIMERA_OK when the x transform handles lines of n points (powers of two up
 * to 8192, other lengths up to 4096 through a power-of-two convolution), else
 * CHIMERA_EUNSUP with chimera_last_error set.  Touches no device.  Lets a
 * Solver refuse its grid at construction; the reference's FFTW plans any
 * length (fb_io.f90:37-38), so it has no counterpart there. */
int32_t chimera_fft_length_check(int32_t n);
/* Form of the mixed-precision (complex64) Hankel and FB-operator GEMMs, for the
 * whole process: 0 = split-BF16, each FP32 operand split exactly into three
 * BF16 pieces and six BF16 MFMA products accumulated in FP32 (default);
 * 1 = FP32 MFMA.  chimera_get_mixed_gemm returns the current form.  No
 * counterpart in the reference, which has no mixed precision. */
int32_t chimera_set_mixed_gemm(int32_t form);
int32_t chimera_get_mixed_gemm(void);
/* Tile autotuner of the Hankel and FB-operator GEMMs (no counterpart in the
 * reference).  on = 1: the first launch of each GEMM shape (element type,
 * form, accumulators, Nx, columns, jobs, K, segments) times every tile shape
 * and keeps the fastest; on = 0 (default): tiles by a block-count rule.  path
 * (may be NULL) names a text file of tuned shapes, read once and appended to.
 * chimera_autotune_entries returns the number of shapes tuned or loaded. */
int32_t chimera_set_autotune(int32_t on, const char* path);
int32_t chimera_autotune_entries(void);
/* Stream (hipStream_t) every later launch goes to; NULL = the null stream. */
int32_t chimera_set_stream(void* hip_stream);
void* chimera_get_stream(void);
int32_t chimera_synchronize(void);

/* Contexts and captured steps (no counterpart in the reference, which runs one
 * simulation per process).  A context owns a stream (default priority, not
 * blocking on the null stream), the scratch memory its launches share and the
 * device copies of the small tables they read.  chimera_context_bind makes a
 * context current for the calling thread (NULL = the default context, whose
 * stream chimera_set_stream sets); every later call of the thread launches on
 * it.  Independent simulations on different contexts run concurrently.
 * chimera_context_query sets *done (HOST) to 1 when all work queued on the
 * context has finished, else 0, without waiting.
 * chimera_capture_begin / _end record the launches of the current context into
 * a graph instead of running them (nothing may allocate, upload a new table or
 * synchronise meanwhile: run the same step once uncaptured first);
 * chimera_graph_launch replays it on its context's stream.  A graph reads the
 * device memory it was recorded with: its arrays must stay allocated and its
 * context alive until chimera_graph_destroy.  chimera_mem_info reports free
 * and total device memory in bytes (HOST outputs). */
int32_t chimera_context_create(void** ctx);
int32_t chimera_context_destroy(void* ctx);
int32_t chimera_context_bind(void* ctx);
void* chimera_context_current(void);
int32_t chimera_context_query(void* ctx, int32_t* done);
int32_t chimera_capture_begin(void);
int32_t chimera_capture_end(void** graph);
int32_t chimera_capture_abort(void);
int32_t chimera_graph_launch(void* graph);
int32_t chimera_graph_destroy(void* graph);
int32_t chimera_device_synchronize(void);
int32_t chimera_mem_info(int64_t* free_bytes, int64_t* total_bytes);

/* Device memory for resident simulations (no counterpart in the reference,
 * whose arrays are NumPy-owned).  chimera_malloc returns zero-filled device
 * memory; chimera_memcpy copies between any two of host/device memory (device
 * to device asynchronously on the library stream, otherwise after draining
 * it); chimera_memset fills bytes with a kernel (safe inside captured graphs);
 * chimera_axpb sets p[i] = a*p[i] + b for n doubles. */
int32_t chimera_malloc(void** ptr, int64_t bytes);
int32_t chimera_free(void* ptr);
int32_t chimera_memcpy(void* dst, const void* src, int64_t bytes);
int32_t chimera_memset(void* ptr, int32_t value, int64_t bytes);
int32_t chimera_axpb(double* p, int64_t n, double a, double b);

/* In-library kernel timing (no counterpart in the reference).  While enabled,
 * the launches of each kernel family are bracketed by HIP events on the
 * library stream; chimera_prof_read sums the elapsed times of one family.
 * Families: 0 position push, 1 current deposition, 2 charge deposition,
 * 3 gather + push, 4 Hankel-transform GEMMs, 5 x FFTs, 6 FB-operator GEMMs,
 * 7 PSATD / elementwise field updates, 8 compaction / sort, 9 other. */
int32_t chimera_prof_enable(int32_t on);
int32_t chimera_prof_reset(void);
int32_t chimera_prof_read(int32_t family, double* total_ms, int64_t* count);

/* ---------------------------------------------------------------- particle_tools.f90 */

/* Relativistic Boris push.  momenta(3,np) in/out, Fld(6,np).
 * Replaces push_velocs, particle_tools.f90:18-56. */
int32_t chimera_push_velocs(double* momenta, const double* Fld, double dt, int32_t np);

/* Leap-frog position push; coord_cntr = half-step positions.
 * Replaces push_coords, particle_tools.f90:58-82. */
int32_t chimera_push_coords(double* coord, const double* momenta, double* coord_cntr, double dt, int32_t np);

/* FixedCell lattice.  coord(4,np) out, *indPart = particles written;
 * RandPackO(nx,nr); PackO complex(PPC).  Replaces genparts, particle_tools.f90:84-128. */
int32_t chimera_genparts(double* coord, int32_t* indPart, const double* Xgrid, const double* Rgrid,
                         const double* RandPackO, const double* PackX, const double* PackR, const double* PackO,
                         int32_t np, int32_t nx, int32_t nr, int32_t PPC);

/* Domain filter: zero-based indices of the particles inside lims, in order.
 * Replaces sortpartsout, particle_tools.f90:130-153. */
int32_t chimera_sortpartsout(int32_t* indx2stay, int32_t* num2stay, const double* coord, const double* lims,
                             int32_t np);

/* Chunk id per particle (-2 outside lims), chunk offsets IndInChnk(0:nchnk)
 * and the count of particles outside.  Xgrid(0:nx).
 * Replaces chunk_coords_boundaries, particle_tools.f90:155-208. */
int32_t chimera_chunk_coords_boundaries(int8_t* chunked_indx, int32_t* IndInChnk, int32_t* GoOut, const double* coord,
                                        const double* lims, const double* Xgrid, int32_t nchnk, int32_t np,
                                        int32_t nx);

/* dat(:,1:np) = dat(:,idx+1) for dat(3,np0) / dat(np0); idx zero-based int64.
 * Replaces align_data_vec / align_data_scl, particle_tools.f90:270-324. */
int32_t chimera_align_data_vec(double* dat, const int64_t* chunked_indx, int32_t np, int32_t np0);
int32_t chimera_align_data_scl(double* dat, const int64_t* chunked_indx, int32_t np, int32_t np0);

/* ---------------------------------------------------------------- grid_deps.f90 / grid_deps_chnk.f90 */

/* Current deposition onto curr(0:nx,0:nr,0:nkO,3) (accumulated), then the
 * ghost-row fold.  Replaces dep_curr, grid_deps.f90:18-87. */
int32_t chimera_dep_curr(const double* coord, const double* momenta, const double* wghts, double* curr, double leftX,
                         const double* Rgrid, double dx_inv, double dr_inv, int32_t np, int32_t nx, int32_t nr,
                         int32_t nkO);

/* Charge deposition onto dens(0:nx,0:nr,0:nkO).  Replaces dep_dens, grid_deps.f90:89-147. */
int32_t chimera_dep_dens(const double* coord, const double* wghts, double* dens, double leftX, const double* Rgrid,
                         double dx_inv, double dr_inv, int32_t np, int32_t nx, int32_t nr, int32_t nkO);

/* Chunked variants: deposit particles IndInChunk(0)..IndInChunk(nchnk)-1 with
 * the unchunked semantics (no guard-cell race, no dropped edge deposits).
 * Replace dep_curr_chnk / dep_dens_chnk, grid_deps_chnk.f90:18-130, :132-234. */
int32_t chimera_dep_curr_chnk(const double* coord, const double* momenta, const double* wghts, double* curr,
                              const int32_t* IndInChunk, int32_t guards, double leftX, const double* Rgrid,
                              double dx_inv, double dr_inv, int32_t np, int32_t nx, int32_t nr, int32_t nkO,
                              int32_t nchnk);
int32_t chimera_dep_dens_chnk(const double* coord, const double* wghts, double* dens, const int32_t* IndInChunk,
                              int32_t guards, double leftX, const double* Rgrid, double dx_inv, double dr_inv,
                              int32_t np, int32_t nx, int32_t nr, int32_t nkO, int32_t nchnk);

/* Field gather: Fld_tot(6,np) += fields of Fld(0:nx,0:nr,0:nkO,6) at the
 * particles.  Replaces proj_fld, grid_deps.f90:149-217. */
int32_t chimera_proj_fld(const double* coord, const double* wghts, const double* Fld, double* Fld_tot, double leftX,
                         const double* Rgrid, double dx_inv, double dr_inv, int32_t np, int32_t nx, int32_t nr,
                         int32_t nkO);

/* Mode normalisation and ghost row of eb_spc(0:nx,0:nr,nkO,6) (nkO = number
 * of modes).  Replaces eb_correction, grid_deps.f90:219-266. */
int32_t chimera_eb_correction(double* eb_spc, int32_t nx, int32_t nr, int32_t nkO);

/* ---------------------------------------------------------------- fb_io.f90 (nkO = number of modes) */

/* Forward FB transform: DHT with In(nr,nkr,nkO) over rows 1..nr of
 * vec(nkx,0:nr,nkO,3), FFT along x, times exp(-i kx leftX) -> vec_fb(nkx,nkr,nkO,3).
 * Replaces fb_vec_in, fb_io.f90:18-59; fb_scl_in (one component), fb_io.f90:61-98. */
int32_t chimera_fb_vec_in(double* vec_fb, const double* vec, double leftX, const double* kx, const double* In,
                          int32_t nkx, int32_t nr, int32_t nkO, int32_t nkr);
int32_t chimera_fb_scl_in(double* scl_fb, const double* scl, double leftX, const double* kx, const double* In,
                          int32_t nkx, int32_t nr, int32_t nkO, int32_t nkr);

/* Inverse FB transform of E (components 1..3 of e_fb(nkx,nkr,nkO,6)) and
 * b_fb(nkx,nkr,nkO,3) into eb_spc(nkx,0:nr,nkO,6) (ghost row zero).
 * Replaces fb_eb_out, fb_io.f90:182-228. */
int32_t chimera_fb_eb_out(double* eb_spc, const double* e_fb, const double* b_fb, double leftX, const double* kx,
                          const double* Out, int32_t nkx, int32_t nr, int32_t nkO, int32_t nkr);

/* Absorbing layer on vec(nkx,nkr,nkO,3); modefilt 0 ('left') only.
 * Replaces fb_filtr, fb_io.f90:230-308. */
int32_t chimera_fb_filtr(double* vec, double leftX, const double* kx, const double* filtr, int32_t modefilt,
                         int32_t nkx, int32_t nkr, int32_t nkO, int32_t nxfilt);

/* ---------------------------------------------------------------- fb_math.f90 (nkO = max mode) */

/* Curl: vec_fb(nkx,nkr,0:nkO,3) -> vec_fb_loc(nkx,nkr_loc,0:nkO,3); DpS2S,
 * DmS2S(nkr,nkr_loc,0:nkO+1).  Replaces fb_rot, fb_math.f90:18-94. */
int32_t chimera_fb_rot(double* vec_fb_loc, const double* vec_fb, const double* DpS2S, const double* DmS2S,
                       const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc);
/* Gradient of scl_fb(nkx,nkr,0:nkO).  Replaces fb_grad, fb_math.f90:96-149. */
int32_t chimera_fb_grad(double* vec_fb_loc, const double* scl_fb, const double* DpS2S, const double* DmS2S,
                        const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc);
/* Divergence.  Replaces fb_div, fb_math.f90:151-199. */
int32_t chimera_fb_div(double* scl_fb_loc, const double* vec_fb, const double* DpS2S, const double* DmS2S,
                       const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc);
/* Gradient of divergence, in place (needs nkr_loc == nkr).
 * Replaces fb_graddiv, fb_math.f90:201-293. */
int32_t chimera_fb_graddiv(double* vec_fb, const double* DpS2S, const double* DmS2S, const double* kx, int32_t nkx,
                           int32_t nkr, int32_t nkO, int32_t nkr_loc);

/* ---------------------------------------------------------------- maxwell_solvers.f90 (nkO = number of modes) */

/* PSATD step with space charge; C1, C2 real (nkx,nkr,nkO,5).
 * Replaces maxwell_push_with_spchrg, maxwell_solvers.f90:18-60. */
int32_t chimera_maxwell_push_with_spchrg(double* EG_fb, const double* j_fb, const double* grad_rho_n_fb,
                                         const double* grad_rho_np1_fb, const double* C1, const double* C2,
                                         int32_t nkx, int32_t nkr, int32_t nkO);
/* PSATD step without space charge; C1, C2 real (nkx,nkr,nkO,3) (the Fortran
 * declares them complex and the driver passes real tables that f2py converts).
 * Replaces maxwell_push_wo_spchrg, maxwell_solvers.f90:62-96. */
int32_t chimera_maxwell_push_wo_spchrg(double* EG_fb, const double* j_fb, const double* C1, const double* C2,
                                       int32_t nkx, int32_t nkr, int32_t nkO);
/* Static initial push; C1, C2 complex (nkx,nkr,nkO,2).
 * Replaces maxwell_init_push, maxwell_solvers.f90:98-129. */
int32_t chimera_maxwell_init_push(double* EG_fb, const double* j_fb, const double* grad_rho_n_fb, const double* C1,
                                  const double* C2, int32_t nkx, int32_t nkr, int32_t nkO);
/* j += (grad div j + (grad rho_np1 - grad rho_n) dt_inv) w2_inv.
 * Replaces poiss_corr, maxwell_solvers.f90:131-164. */
int32_t chimera_poiss_corr(double* j_fb, const double* grad_div_j_fb, const double* grad_rho_n_fb,
                           const double* grad_rho_np1_fb, double dt_inv, const double* w2_inv, int32_t nkx,
                           int32_t nkr, int32_t nkO);
/* j += (grad div j + DT grad rho_n) w2_inv with DT(nkx) complex (the static
 * kick's Poisson correction).  Replaces poiss_corr_stat, maxwell_solvers.f90:166-197. */
int32_t chimera_poiss_corr_stat(double* j_fb, const double* grad_div_j_fb, const double* grad_rho_n_fb, const double* DT,
                                const double* w2_inv, int32_t nkx, int32_t nkr, int32_t nkO);
/* EG_fb(nkx,nkr,nkO,6) *= exp(-0.5 i dt beta0 kx) (free drift over one step).
 * Replaces field_drift, maxwell_solvers.f90:199-226. */
int32_t chimera_field_drift(double* EG_fb, const double* kx, double beta0, double dt, int32_t nkx, int32_t nkr,
                            int32_t nkO);
/* Elementwise helpers.  Replace omp_mult_vec/scl, omp_add_vec/scl, maxwell_solvers.f90:228-318. */
int32_t chimera_omp_mult_vec(double* vec_fb, const double* A, int32_t nkx, int32_t nkr, int32_t nkO);
int32_t chimera_omp_mult_scl(double* scl_fb, const double* A, int32_t nkx, int32_t nkr, int32_t nkO);
int32_t chimera_omp_add_vec(double* vec_fb, const double* A, int32_t nkx, int32_t nkr, int32_t nkO);
int32_t chimera_omp_add_scl(double* scl_fb, const double* A, int32_t nkx, int32_t nkr, int32_t nkO);

/* ---------------------------------------------------------------- devices.f90 */

/* Analytic planar undulator, params = (a0, lambda, X0, Lx); Fld(6,np) += field.
 * Replaces Undul_Analytic, devices.f90:162-203. */
int32_t chimera_undul_analytic(const double* coord, double* Fld, double t, const double* params, int32_t np);
/* Linearly tapered analytic undulator, params = (a0, lambda, X0, Lx, taper): the
 * amplitude of chimera_undul_analytic times 1 + taper (x - X0 - Lx/2) / (Lx/2).
 * Replaces Undul_Analytic_Taper, devices.f90:117-160. */
int32_t chimera_undul_analytic_taper(const double* coord, double* Fld, double t, const double* params, int32_t np);
/* Mapped undulator: a0(2,nx) on the nodes Xleft + i dx (i = 0..nx-1), params =
 * (lambda, Xleft, dx); Fld(5) += (quadratic interpolation of a0(1,:)) cosh(ku y),
 * Fld(4) += (that of a0(2,:)) sinh(ku y) for x in [Xleft + dx, Xleft + (nx-1) dx].
 * A node outside the map contributes zero.  Replaces Undul_Mapped, devices.f90:18-62. */
int32_t chimera_undul_mapped(const double* coord, double* Fld, double t, const double* a0, const double* params,
                             int32_t np, int32_t nx);
/* The same, tapered: params = (lambda, Xleft, dx, Lx, taper), amplitude times
 * 1 + taper (x - x_shift - Xleft - Lx/2) / (Lx/2), x_shift = (nx dx - Lx)/2.
 * Replaces Undul_Mapped_Tap, devices.f90:64-115. */
int32_t chimera_undul_mapped_tap(const double* coord, double* Fld, double t, const double* a0, const double* params,
                                 int32_t np, int32_t nx);
/* Plane wave, params = (a0, lambda, X0, Lx, ramp, theta, phi0): with the ramped
 * amplitude A(x), E = a0 A sin(k0 (x cos theta + y sin theta - t) + phi0);
 * Fld(3) += E, Fld(4) += E sin theta, Fld(5) -= E cos theta.
 * Replaces PlaneWave, devices.f90:205-252. */
int32_t chimera_planewave(const double* coord, double* Fld, double t, const double* params, int32_t np);
/* Gaussian beam, params = (lambda, axis, x0, y0, z0, Lx, Ly, Lz):
 * E = a0 exp(-xp^2/Lx^2 - yp^2/Ly^2 - zp^2/Lz^2) sin(2 pi xp / lambda) with
 * xp = x - x0 - axis time, yp = y - y0, zp = z - z0; Fld(3) += E, Fld(5) -= axis E.
 * Replaces GaussBeam, devices.f90:254-297. */
int32_t chimera_gaussbeam(const double* coord, double* Fld, double time, double a0, const double* params, int32_t np);

/* ---------------------------------------------------------------- envelope (KxShift) solver
 * The f2py entry points of grid_deps_env.f90, grid_deps_env_chnk.f90 and
 * fb_math_env.f90.  Envelope grids carry the modes -nkO..nkO (mode axis of
 * length 2 nkO + 1, Fortran order, e.g. curr(0:nx,0:nr,-nkO:nkO,3)); the
 * coupling matrices DpS2S/DmS2S (nkr, nkr_loc, -nkO-1:nkO+1).  Arrays host or
 * device memory, as the drop-in tier above. */
/* = dep_curr_env (grid_deps_env.f90:18-95): the third current component only,
 * complex weight w e^{-i kx0 x}, then the ghost fold. */
int32_t chimera_dep_curr_env(const double* coord, const double* momenta, const double* wghts, double* curr,
                             double leftX, const double* Rgrid, double dx_inv, double dr_inv, double kx0, int32_t np,
                             int32_t nx, int32_t nr, int32_t nkO);
/* = dep_dens_env (grid_deps_env.f90:97-162): the complex weight applied twice. */
int32_t chimera_dep_dens_env(const double* coord, const double* wghts, double* dens, double leftX, const double* Rgrid,
                             double dx_inv, double dr_inv, double kx0, int32_t np, int32_t nx, int32_t nr,
                             int32_t nkO);
/* = dep_curr_env_chnk / dep_dens_env_chnk (grid_deps_env_chnk.f90:18-261): the
 * particles IndInChunk(0)..IndInChunk(nchnk)-1 deposited with the unchunked
 * semantics (the reference's guard merges drop or race on chunk edges). */
int32_t chimera_dep_curr_env_chnk(const double* coord, const double* momenta, const double* wghts, double* curr,
                                  const int32_t* IndInChunk, int32_t guards, double leftX, const double* Rgrid,
                                  double dx_inv, double dr_inv, double kx0, int32_t np, int32_t nx, int32_t nr,
                                  int32_t nkO, int32_t nchnk);
int32_t chimera_dep_dens_env_chnk(const double* coord, const double* wghts, double* dens, const int32_t* IndInChunk,
                                  int32_t guards, double leftX, const double* Rgrid, double dx_inv, double dr_inv,
                                  double kx0, int32_t np, int32_t nx, int32_t nr, int32_t nkO, int32_t nchnk);
/* = proj_fld_env (grid_deps_env.f90:164-238); Fld(0:nx,0:nr,-nkO:nkO,6). */
int32_t chimera_proj_fld_env(const double* coord, const double* wghts, const double* Fld, double* Fld_tot,
                             double leftX, const double* Rgrid, double dx_inv, double dr_inv, double kx0, int32_t np,
                             int32_t nx, int32_t nr, int32_t nkO);
/* = eb_correction_env (grid_deps_env.f90:240-283). */
int32_t chimera_eb_correction_env(double* eb_spc, int32_t nx, int32_t nr, int32_t nkO);
/* = fb_rot_env, fb_grad_env, fb_div_env, fb_graddiv_env (fb_math_env.f90:18-233). */
int32_t chimera_fb_rot_env(double* vec_fb_loc, const double* vec_fb, const double* DpS2S, const double* DmS2S,
                           const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc);
int32_t chimera_fb_grad_env(double* vec_fb_loc, const double* scl_fb, const double* DpS2S, const double* DmS2S,
                            const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc);
int32_t chimera_fb_div_env(double* scl_fb_loc, const double* vec_fb, const double* DpS2S, const double* DmS2S,
                           const double* kx, int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc);
int32_t chimera_fb_graddiv_env(double* vec_fb, const double* DpS2S, const double* DmS2S, const double* kx,
                               int32_t nkx, int32_t nkr, int32_t nkO, int32_t nkr_loc);
/* = maxwell_push_wo_spchrg (maxwell_solvers.f90:62-96) with the complex tables
 * C1, C2 (nkx,nkr,nkO,3) complex128 of the envelope solver (solvers.py:244-267);
 * nkO is the mode count here, as in maxwell_solvers.f90. */
int32_t chimera_maxwell_push_wo_spchrg_env(double* EG_fb, const double* j_fb, const double* C1, const double* C2,
                                           int32_t nkx, int32_t nkr, int32_t nkO);

/* ================================================================ resident tier
 *
 * Entry points for simulations whose state stays in device memory (the
 * chimera_amd.Solver / Specie / ChimeraRun driver).  Every array pointer is
 * device memory and the calls are asynchronous on the library stream, except
 * where a host output is named.  Particles are structure-of-arrays: component
 * c of particle i at coord[c*ld + i].  Grid arrays keep the reference layout.
 * These fuse steps that the reference runs as separate f2py calls; each cites
 * the calls it stands for.
 */

/* Leap-frog push, SoA.  = push_coords (particle_tools.f90:58-82). */
int32_t chimera_soa_push_coords(double* coord, double* coord_half, const double* momenta, int64_t ld, int32_t np,
                                double dt);

/* Deposition WITHOUT the ghost fold (chimera_fb_forward applies it with fold=1).
 * r0 = Rgrid(0), rmax = Rgrid(nr).  = dep_curr / dep_dens particle loops
 * (grid_deps.f90:34-78, :104-142). */
int32_t chimera_soa_dep_curr(const double* coord, const double* momenta, const double* wghts, int64_t ld, int32_t np,
                             double* curr, double leftX, const double* Rgrid, double dx_inv, double dr_inv, double r0,
                             double rmax, int32_t nx, int32_t nr, int32_t nkO);
int32_t chimera_soa_dep_dens(const double* coord, const double* wghts, int64_t ld, int32_t np, double* dens,
                             double leftX, const double* Rgrid, double dx_inv, double dr_inv, double r0, double rmax,
                             int32_t nx, int32_t nr, int32_t nkO);

/* get_dens_on_grid (species.py:400-427): dens(nx+1, nr+1, nkO+1) complex128 =
 * dep_dens of the particles (ghost fold included, grid_deps.f90:89-147) times
 * vgrid[ir] = 1/(2 pi dx dr Rgrid[ir]) (0 where Rgrid <= 0) per radial row.
 * dens and vgrid in device memory; dens is overwritten.  The _mixed form
 * reads float weights (the deposit and dens stay FP64). */
int32_t chimera_soa_dens_on_grid(const double* coord, const double* wghts, int64_t ld, int32_t np, double* dens,
                                 double leftX, const double* Rgrid, double dx_inv, double dr_inv, double r0,
                                 double rmax, int32_t nx, int32_t nr, int32_t nkO, const double* vgrid);
int32_t chimera_soa_dens_on_grid_mixed(const double* coord, const float* wghts, int64_t ld, int32_t np,
                                       double* dens, double leftX, const double* Rgrid, double dx_inv, double dr_inv,
                                       double r0, double rmax, int32_t nx, int32_t nr, int32_t nkO,
                                       const double* vgrid);

/* Window frames on the device.  The resident entry points below that take a
 * window edge (leftX) also take `frame`, a DEVICE pointer or NULL: when it is
 * not NULL the kernels read the edge from it at run time and ignore the
 * by-value argument, so that a step captured into a graph
 * (chimera_capture_begin) replays correctly while the window moves.  A frame
 * is two doubles: frame[0] = the edge for the stage-1 consumers of a step
 * (J deposit and its forward transform), frame[1] = the edge after the staged
 * stage-2 move (rho deposit, its transform, the inverse transform, gather).
 * chimera_frame_step advances it as move_frame does (chimera_main.py:286-291,
 * Xgrid + shiftX): frame[0] = frame[1] + shift1, then frame[1] = frame[0] +
 * shift2 (a shift of 0 is no move), in FP64. */
int32_t chimera_frame_step(double* frame, double shift1, double shift2);

/* Fused position push and deposition: with dt != 0 the positions are pushed
 * first (coord -> coord at n+1, coord_half = midpoint), otherwise coord_half is
 * read as given; then J (curr, if not NULL) is deposited from coord_half and
 * momenta with leftX_J, and rho (dens, if not NULL) from coord with leftX_rho
 * (frame[0] and frame[1] instead when frame is not NULL).
 * Deposits are accumulated per block in LDS tiles and added to the grid
 * unfolded.  = push_coords + dep_curr + dep_dens (particle_tools.f90:58-82,
 * grid_deps.f90:18-147), as make_step runs them (chimera_main.py:84-88). */
int32_t chimera_soa_push_deposit(double* coord, double* coord_half, const double* momenta, const double* wghts,
                                 int64_t ld, int32_t np, double dt, double leftX_J, double leftX_rho,
                                 const double* Rgrid, double dx_inv, double dr_inv, double r0, double rmax, int32_t nx,
                                 int32_t nr, int32_t nkO, double* curr, double* dens, const double* frame);

/* Gather of Fld(0:nx,0:nr,0:nkO,6) at the particles, plus the analytic
 * undulator when undul_params (HOST double[4]) is not NULL, then the Boris push
 * with dt = push_dt.  = make_field + proj_fld + undul_analytic + push_velocs
 * (species.py:246-298, grid_deps.f90:149-217, devices.f90:162-203,
 * particle_tools.f90:18-56). */
int32_t chimera_soa_gather_push(const double* coord, double* momenta, const double* wghts, int64_t ld, int32_t np,
                                const double* Fld, double leftX, const double* Rgrid, double dx_inv, double dr_inv,
                                double r0, double rmax, int32_t nx, int32_t nr, int32_t nkO, double push_dt,
                                const double* undul_params, const double* frame);

/* An external-field device of devices.f90, as a species' Devices entry
 * (species.py:258-277) describes it: kind 0 undul_analytic (params a0, lambda,
 * X0, Lx), 1 undul_analytic_taper (+ taper), 2 undul_mapped (lambda, Xleft,
 * dx; map), 3 undul_mapped_tap (+ Lx, taper), 4 planewave (a0, lambda, X0,
 * Lx, ramp, theta, phi0), 5 gaussbeam (lambda, axis, x0, y0, z0, Lx, Ly, Lz;
 * amplitude a0).  map: the field map a0(2, nmap) in DEVICE memory. */
typedef struct {
    int32_t kind;
    int32_t nmap;
    const double* map;
    double a0;
    double params[8];
} chimera_device_t;
/* chimera_soa_gather_push with up to 4 devices (HOST array of ndev
 * descriptors) added in order to the gathered field at time t (i_step *
 * TimeStep), before the push.  = make_field + proj_fld + make_device +
 * push_velocs (species.py:246-298, devices.f90:18-297). */
int32_t chimera_soa_gather_push_dev(const double* coord, double* momenta, const double* wghts, int64_t ld,
                                    int32_t np, const double* Fld, double leftX, const double* Rgrid, double dx_inv,
                                    double dr_inv, double r0, double rmax, int32_t nx, int32_t nr, int32_t nkO,
                                    double push_dt, const chimera_device_t* devices, int32_t ndev, double t,
                                    const double* frame);
/* Resident tier of the envelope path: structure-of-arrays particles (component c
 * of particle i at c*ld + i), every array in device memory, asynchronous.
 * = dep_curr_env / dep_dens_env WITHOUT the ghost fold (chimera_fb_forward
 * with fold=1 applies it), and proj_fld_env fused with the species' devices
 * and push_velocs (chimera_main.py:130-151, :91-92; devices as for
 * chimera_soa_gather_push_dev).  Rgrid(0) = r0 and Rgrid(nr) = rmax come from
 * the host; frame as for the other resident calls. */
int32_t chimera_soa_dep_curr_env(const double* coord_half, const double* momenta, const double* wghts, int64_t ld,
                                 int32_t np, double* curr, double leftX, const double* Rgrid, double dx_inv,
                                 double dr_inv, double r0, double rmax, double kx0, int32_t nx, int32_t nr,
                                 int32_t nkO, const double* frame);
int32_t chimera_soa_dep_dens_env(const double* coord, const double* wghts, int64_t ld, int32_t np, double* dens,
                                 double leftX, const double* Rgrid, double dx_inv, double dr_inv, double r0,
                                 double rmax, double kx0, int32_t nx, int32_t nr, int32_t nkO, const double* frame);
int32_t chimera_soa_gather_push_env(const double* coord, double* momenta, const double* wghts, int64_t ld, int32_t np,
                                    const double* Fld, double leftX, const double* Rgrid, double dx_inv,
                                    double dr_inv, double r0, double rmax, double kx0, int32_t nx, int32_t nr,
                                    int32_t nkO, double push_dt, const chimera_device_t* devices, int32_t ndev,
                                    double t, const double* frame);

/* Stable removal of the particles outside lims (HOST double[4]: x range,
 * r^2 range), tested on coord (test_half=0) or coord_half (1); *np_out (HOST)
 * receives the new count.  = sortpartsout + align_data_* (species.py:385-398). */
int32_t chimera_soa_compact(double* coord, double* coord_half, double* momenta, double* wghts, int64_t ld, int32_t np,
                            const double* lims, int32_t test_half, int32_t* np_out);
/* The same selection written to a second set of arrays (leading dimension
 * ld_dst >= np) and the sources left as they were: a caller that keeps spare
 * arrays swaps them in instead of copying the survivors back. */
int32_t chimera_soa_compact_into(const double* coord, const double* coord_half, const double* momenta,
                                 const double* wghts, int64_t ld, int32_t np, double* d_coord, double* d_half,
                                 double* d_mom, double* d_w, int64_t ld_dst, const double* lims, int32_t test_half,
                                 int32_t* np_out);

/* Stable counting sort of the particles by cell, absorbing those outside lims
 * (HOST double[4], tested on coord or, with test_half, coord_half; NULL keeps
 * all).  Cell key = ir*nx + ix from the expressions of grid_deps.f90:46-47 with
 * nx, nr the CELL counts, kept particles outside the grid after all cells.
 * The sorted, compacted particles are written to the second set of arrays
 * (leading dimension ld2); *np_out (HOST) receives their number and perm
 * (device, np entries, may be NULL) the source index of each.
 * = chunk_coords_boundaries + argsort + align_data_* (species.py:378-398),
 * with cells as the chunks and a stable order. */
int32_t chimera_soa_sort_cells(const double* coord, const double* coord_half, const double* momenta,
                               const double* wghts, int64_t ld, int32_t np, double* coord2, double* coord_half2,
                               double* momenta2, double* wghts2, int64_t ld2, double leftX, double dx_inv,
                               double dr_inv, double r0, int32_t nx, int32_t nr, const double* lims, int32_t test_half,
                               int32_t* np_out, int32_t* perm);
/* The same with float32 momenta and weights (the species of a mixed-precision run). */
int32_t chimera_soa_sort_cells_mixed(const double* coord, const double* coord_half, const float* momenta,
                                     const float* wghts, int64_t ld, int32_t np, double* coord2, double* coord_half2,
                                     float* momenta2, float* wghts2, int64_t ld2, double leftX, double dx_inv,
                                     double dr_inv, double r0, int32_t nx, int32_t nr, const double* lims,
                                     int32_t test_half, int32_t* np_out, int32_t* perm);

/* Forward FB transform of ncomp components, fb_vec_in/fb_scl_in with nkO =
 * nmodes, times scale(nkx,nkr,nmodes) (DepFact) when not NULL.  With fold=1
 * the deposition ghost fold is first applied to `in` in place (row 1 -= row 0,
 * row 0 = 0), which leaves `in` as the reference's dep_curr/dep_dens return it.
 * = dep fold + fb_vec_in + omp_mult_vec (grid_deps.f90:80-85, fb_io.f90:18-98,
 * solvers.py:417-435). */
int32_t chimera_fb_forward(double* out, double* in, double leftX, const double* kx, const double* In,
                           const double* scale, int32_t nkx, int32_t nr, int32_t nmodes, int32_t nkr, int32_t ncomp,
                           int32_t fold, const double* frame);

/* Inverse FB transform into eb_spc(nkx,0:nr,nmodes,6); ghost=1: ghost row zero
 * (fb_eb_out alone), ghost=2: eb_correction applied as well, ghost=3 or 4:
 * eb_correction_env (grid_deps_env.f90:240-283: every mode times 1/pi, the
 * ghost row +row 1 (3, the reference's choice for nkO = 0) or -row 1 (4)).
 * = fb_eb_out + eb_correction (fb_io.f90:182-228, grid_deps.f90:219-266). */
int32_t chimera_fb_inverse(double* eb_spc, const double* e_fb, const double* b_fb, double leftX, const double* kx,
                           const double* Out, int32_t nkx, int32_t nr, int32_t nmodes, int32_t nkr, int32_t ghost,
                           const double* frame);

/* B = rot(G) * scale (PoissFact).  = fb_rot + omp_mult_vec (solvers.py:536-553). */
int32_t chimera_fb_rot_scaled(double* vec_fb_loc, const double* vec_fb, const double* DpS2S, const double* DmS2S,
                              const double* kx, const double* scale, int32_t nkx, int32_t nkr, int32_t nkO);

/* One Poisson-correction iteration, j += (graddiv j + (gp - gn) dt_inv) w2_inv.
 * = fb_graddiv + poiss_corr (solvers.py:317-326). */
int32_t chimera_fb_graddiv_poisson(double* j_fb, const double* grad_rho_n_fb, const double* grad_rho_np1_fb,
                                   double dt_inv, const double* w2_inv, const double* DpS2S, const double* DmS2S,
                                   const double* kx, int32_t nkx, int32_t nkr, int32_t nkO);

/* op = 0 rot, 1 grad, 2 div, 3 graddiv (in may equal out only for 3), nkr_loc = nkr.
 * = fb_rot / fb_grad / fb_div / fb_graddiv (fb_math.f90). */
int32_t chimera_fb_op(int32_t op, double* out, const double* in, const double* DpS2S, const double* DmS2S,
                      const double* kx, int32_t nkx, int32_t nkr, int32_t nkO);

/* Absorbing layer on ncomp components of vec(nkx,nkr,nmodes,ncomp).
 * = fb_filtr with modefilt 0 (fb_io.f90:230-270). */
int32_t chimera_fb_filter(double* vec, double leftX, const double* kx, const double* filtr, int32_t nxfilt,
                          int32_t nkx, int32_t nkr, int32_t nmodes, int32_t ncomp);

/* ======================================================================
 * Diagnostics on the device (the reference computes them on the host from
 * the particle arrays).
 * ====================================================================== */
/* The 13 weighted sums behind the beam envelopes of get_beam_envelops
 * (diagnostics.py:177-207): sums[0..12] = sum w, sum x w, sum y w, sum z w,
 * sum x^2 w, sum y^2 w, sum z^2 w, sum px^2 w, sum py^2 w, sum pz^2 w,
 * sum x px w, sum y py w, sum z pz w over the np particles (x, y, z the rows
 * of coord_half, px, py, pz those of momenta, component stride ld).  The
 * particle arrays are in device memory; sums in device memory (the call is
 * asynchronous) or host memory (the call waits).  FP64 sums in an order fixed
 * by np, so repeated calls give the same bits. */
int32_t chimera_beam_moments(const double* coord_half, const double* momenta, const double* wghts, int64_t ld,
                             int32_t np, double* sums);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  chimera_beam_moments with the same kernels and the same
 * bits, but its partial sums live in a fixed 53,248-byte device array of the
 * library (512 workgroups x 13 doubles, any np) and sums must be device
 * memory (13 doubles): nothing is allocated, staged or freed.  Asynchronous on
 * the library stream.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU race. */
int32_t chimera_beam_moments_noalloc(const double* coord_half, const double* momenta, const double* wghts,
                                     int64_t ld, int32_t np, double* sums);
/* px0 = sum(px w) / sum(w) over the np particles (px the first row of
 * momenta), the mean momentum of the StaticKick field update, FP64 sums in an
 * order fixed by np; NaN for np = 0.  Replaces the PXmean expression of
 * update_fields, chimera_main.py:120-121.  Device particle arrays; px0 in
 * device or host memory, as for chimera_beam_moments. */
int32_t chimera_momentum_mean(const double* momenta, const double* wghts, int32_t np, double* px0);

/* The particles phs_out dumps (diagnostics.py:89-107): those with
 * sum_c (1 + u_c^2) > 20^2, in their order, as PartsPack(7, n) (rows: x, y, z
 * at the half step, px, py, pz, w; Fortran order, FP64).  Device particle
 * arrays of component stride ld.  out = NULL only counts: *n_sel (host) = n.
 * Otherwise out (device, room for cap particles) receives them; the call
 * waits for the kernels. */
int32_t chimera_phs_select(const double* coord_half, const double* momenta, const double* wghts, int64_t ld,
                           int32_t np, double* out, int64_t cap, int64_t* n_sel);
int32_t chimera_phs_select_mixed(const double* coord_half, const float* momenta, const float* wghts, int64_t ld,
                                 int32_t np, double* out, int64_t cap, int64_t* n_sel);
/* intens_profO (utils.f90:18-57; f2py name intens_profo): PWR_RO(NO, nr) =
 * sum over l < 3 and ix of |sum_m phaseO(m) Fld(ix, ir, m, l)|^2, ir = 1..nr,
 * for Fld(nkx, 0:nr, -nkO:nkO, 3) (nmodes = 2 nkO + 1, odd, as the
 * reference's dimension requires) and phaseO(m) = exp(i m theta), theta =
 * 2 pi (iO-1)/(NO-1) formed by the reference's recurrences.  Fld and PWR_RO
 * host or device. */
int32_t chimera_intens_profo(double* PWR_RO, const double* Fld, int32_t NO, int32_t nkx, int32_t nr, int32_t nmodes);
int32_t chimera_intens_profo_c8(double* PWR_RO, const float* Fld, int32_t NO, int32_t nkx, int32_t nr,
                                int32_t nmodes);
/* EM energy per kx of nrg_out (diagnostics.py:109-122), before its reordering
 * of the kx axis: out[x] = sum_{k,m} EnergyFact[x,k,m] * sum_{c<3} |EG_fb[x,k,m,c]|^2
 * with EG_fb(nkx,nkr,nmodes,6) and EnergyFact(nkx,nkr,nmodes) (solvers.py:142-149)
 * in device memory; out (nkx doubles) in device memory (asynchronous) or host
 * memory (the call waits).  FP64 sums in a fixed order. */
int32_t chimera_field_energy(const double* EG_fb, const double* energy_fact, int32_t nkx, int32_t nkr, int32_t nmodes,
                             double* out);

/* Laser power per x of pwr_out (diagnostics.py:124-149), before its factor
 * PwrFctr*2*dr: out[x] = sum_{r=0..nr} Rgrid[r] * sum_m sum_{c<3} |vec[x,r,m,c]|^2
 * for vec(nkx,0:nr,nmodes,>=3) in device memory (the field of fb_vec_out,
 * fb_io.f90:100-140, or of chimera_fb_inverse with ghost = 1) and Rgrid(0:nr)
 * in device memory; out (nkx doubles) in device or host memory. */
int32_t chimera_line_power(const double* vec, const double* Rgrid, int32_t nkx, int32_t nr, int32_t nmodes,
                           double* out);

/* ======================================================================
 * Mixed-precision resident tier (no counterpart in the reference, which is
 * FP64 throughout).  Particle positions stay FP64; momenta and weights are
 * float32; every grid and FB-space array is complex64 (interleaved float
 * pairs) with the same Fortran layout as the FP64 entry points above; the
 * Hankel, D+- and PSATD tables are the FP64 tables rounded to float32; kx
 * stays FP64.  The GEMMs run on FP32 MFMA with FP32 accumulation.  Each
 * function has the argument list and meaning of the FP64 function of the same
 * name without the suffix.
 * ====================================================================== */
int32_t chimera_soa_push_deposit_mixed(double* coord, double* coord_half, const float* momenta, const float* wghts,
                                       int64_t ld, int32_t np, double dt, double leftX_J, double leftX_rho,
                                       const double* Rgrid, double dx_inv, double dr_inv, double r0, double rmax,
                                       int32_t nx, int32_t nr, int32_t nkO, float* curr, float* dens,
                                       const double* frame);
int32_t chimera_soa_gather_push_mixed(const double* coord, float* momenta, const float* wghts, int64_t ld, int32_t np,
                                      const float* Fld, double leftX, const double* Rgrid, double dx_inv,
                                      double dr_inv, double r0, double rmax, int32_t nx, int32_t nr, int32_t nkO,
                                      double push_dt, const double* undul_params, const double* frame);
int32_t chimera_soa_gather_push_dev_mixed(const double* coord, float* momenta, const float* wghts, int64_t ld,
                                          int32_t np, const float* Fld, double leftX, const double* Rgrid,
                                          double dx_inv, double dr_inv, double r0, double rmax, int32_t nx, int32_t nr,
                                          int32_t nkO, double push_dt, const chimera_device_t* devices, int32_t ndev,
                                          double t, const double* frame);
int32_t chimera_soa_compact_mixed(double* coord, double* coord_half, float* momenta, float* wghts, int64_t ld,
                                  int32_t np, const double* lims, int32_t test_half, int32_t* np_out);
int32_t chimera_soa_compact_into_mixed(const double* coord, const double* coord_half, const float* momenta,
                                       const float* wghts, int64_t ld, int32_t np, double* d_coord, double* d_half,
                                       float* d_mom, float* d_w, int64_t ld_dst, const double* lims,
                                       int32_t test_half, int32_t* np_out);
int32_t chimera_beam_moments_mixed(const double* coord_half, const float* momenta, const float* wghts, int64_t ld,
                                   int32_t np, double* sums);
int32_t chimera_momentum_mean_mixed(const float* momenta, const float* wghts, int32_t np, double* px0);
int32_t chimera_field_energy_c8(const float* EG_fb, const double* energy_fact, int32_t nkx, int32_t nkr,
                                int32_t nmodes, double* out);
int32_t chimera_line_power_c8(const float* vec, const double* Rgrid, int32_t nkx, int32_t nr, int32_t nmodes,
                              double* out);
int32_t chimera_fb_forward_c8(float* out, float* in, double leftX, const double* kx, const float* In,
                              const float* scale, int32_t nkx, int32_t nr, int32_t nmodes, int32_t nkr, int32_t ncomp,
                              int32_t fold, const double* frame);
int32_t chimera_fb_inverse_c8(float* eb_spc, const float* e_fb, const float* b_fb, double leftX, const double* kx,
                              const float* Out, int32_t nkx, int32_t nr, int32_t nmodes, int32_t nkr, int32_t ghost,
                              const double* frame);
int32_t chimera_fb_rot_scaled_c8(float* vec_fb_loc, const float* vec_fb, const float* DpS2S, const float* DmS2S,
                                 const double* kx, const float* scale, int32_t nkx, int32_t nkr, int32_t nkO);
int32_t chimera_fb_graddiv_poisson_c8(float* j_fb, const float* grad_rho_n_fb, const float* grad_rho_np1_fb,
                                      double dt_inv, const float* w2_inv, const float* DpS2S, const float* DmS2S,
                                      const double* kx, int32_t nkx, int32_t nkr, int32_t nkO);
int32_t chimera_fb_op_c8(int32_t op, float* out, const float* in, const float* DpS2S, const float* DmS2S,
                         const double* kx, int32_t nkx, int32_t nkr, int32_t nkO);
int32_t chimera_fb_filter_c8(float* vec, double leftX, const double* kx, const float* filtr, int32_t nxfilt,
                             int32_t nkx, int32_t nkr, int32_t nmodes, int32_t ncomp);
int32_t chimera_poiss_corr_stat_c8(float* j_fb, const float* grad_div_j_fb, const float* grad_rho_n_fb,
                                   const double* DT, const float* w2_inv, int32_t nkx, int32_t nkr, int32_t nkO);
int32_t chimera_field_drift_c8(float* EG_fb, const double* kx, double beta0, double dt, int32_t nkx, int32_t nkr,
                               int32_t nkO);
int32_t chimera_maxwell_push_with_spchrg_c8(float* EG_fb, const float* j_fb, const float* grad_rho_n_fb,
                                            const float* grad_rho_np1_fb, const float* C1, const float* C2,
                                            int32_t nkx, int32_t nkr, int32_t nkO);
int32_t chimera_maxwell_push_wo_spchrg_c8(float* EG_fb, const float* j_fb, const float* C1, const float* C2,
                                          int32_t nkx, int32_t nkr, int32_t nkO);
int32_t chimera_maxwell_init_push_c8(float* EG_fb, const float* j_fb, const float* grad_rho_n_fb, const float* C1,
                                     const float* C2, int32_t nkx, int32_t nkr, int32_t nkO);
int32_t chimera_omp_mult_vec_c8(float* vec_fb, const float* A, int32_t nkx, int32_t nkr, int32_t nkO);
int32_t chimera_omp_add_vec_c8(float* vec_fb, const float* A, int32_t nkx, int32_t nkr, int32_t nkO);

#ifdef __cplusplus
}
#endif

#endif /* CHIMERA_HIP_H */
