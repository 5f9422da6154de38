// fields.hip -- elementwise field kernels of maxwell_solvers.f90: the PSATD
// update with and without space charge, the static initial push, the Poisson
// correction of the current and the multiply/add helpers.  All arrays are
// (nkx, nkr, nkO, ncomp) complex with nkO the number of modes; element e =
// x + nkx*(k + nkr*m) and component stride n = nkx*nkr*nkO.
#include <initializer_list>

#include "common.hpp"
#include "../../include/chimera_hip.h"

namespace chimera {

// maxwell_solvers.f90:18-60
template <typename R>
__global__ void k_psatd_spchrg(cplx_t<R>* __restrict__ EG, const cplx_t<R>* __restrict__ J,
                               const cplx_t<R>* __restrict__ Gn, const cplx_t<R>* __restrict__ Gp,
                               const R* __restrict__ C1, const R* __restrict__ C2, int64_t n) {
    typedef cplx_t<R> C_;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const R a1 = C1[e], a2 = C1[e + n], a3 = C1[e + 2 * n], a4 = C1[e + 3 * n], a5 = C1[e + 4 * n];
    const R b1 = C2[e], b2 = C2[e + n], b3 = C2[e + 2 * n], b4 = C2[e + 3 * n], b5 = C2[e + 4 * n];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        const C_ E = EG[e + l * n], G = EG[e + (3 + l) * n];
        const C_ j = J[e + l * n], gn = Gn[e + l * n], gp = Gp[e + l * n];
        EG[e + l * n] = cadd(cadd(cadd(cadd(crs(a1, E), crs(a2, G)), crs(a3, j)), crs(a4, gn)), crs(a5, gp));
        EG[e + (3 + l) * n] = cadd(cadd(cadd(cadd(crs(b1, E), crs(b2, G)), crs(b3, j)), crs(b4, gn)), crs(b5, gp));
    }
}

// maxwell_solvers.f90:62-96 (real tables)
template <typename R>
__global__ void k_psatd_wo(cplx_t<R>* __restrict__ EG, const cplx_t<R>* __restrict__ J, const R* __restrict__ C1,
                           const R* __restrict__ C2, int64_t n) {
    typedef cplx_t<R> C_;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const R a1 = C1[e], a2 = C1[e + n], a3 = C1[e + 2 * n];
    const R b1 = C2[e], b2 = C2[e + n], b3 = C2[e + 2 * n];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        const C_ E = EG[e + l * n], G = EG[e + (3 + l) * n], j = J[e + l * n];
        EG[e + l * n] = cadd(cadd(crs(a1, E), crs(a2, G)), crs(a3, j));
        EG[e + (3 + l) * n] = cadd(cadd(crs(b1, E), crs(b2, G)), crs(b3, j));
    }
}

// maxwell_solvers.f90:62-96 with the complex tables of the envelope solver
// (solvers.py:244-267): EG' = C1.(E, G, J), G' = C2.(E, G, J) per element.
template <typename R>
__global__ void k_psatd_wo_cplx(cplx_t<R>* __restrict__ EG, const cplx_t<R>* __restrict__ J,
                                const cplx_t<R>* __restrict__ C1, const cplx_t<R>* __restrict__ C2, int64_t n) {
    typedef cplx_t<R> C_;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const C_ a1 = C1[e], a2 = C1[e + n], a3 = C1[e + 2 * n];
    const C_ b1 = C2[e], b2 = C2[e + n], b3 = C2[e + 2 * n];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        const C_ E = EG[e + l * n], G = EG[e + (3 + l) * n], j = J[e + l * n];
        EG[e + l * n] = cadd(cadd(cmul(a1, E), cmul(a2, G)), cmul(a3, j));
        EG[e + (3 + l) * n] = cadd(cadd(cmul(b1, E), cmul(b2, G)), cmul(b3, j));
    }
}

template <typename R>
int32_t psatd_wo_cplx(cplx_t<R>* EG, const cplx_t<R>* J, const cplx_t<R>* C1, const cplx_t<R>* C2, int64_t n) {
    KTimer kt(KF_PSATD);
    hipLaunchKernelGGL((k_psatd_wo_cplx<R>), blocks_for(n, 256), 256, 0, stream(), EG, J, C1, C2, n);
    CH_LAUNCH_CHECK("k_psatd_wo_cplx");
    return ST_OK;
}

// maxwell_solvers.f90:98-129 (complex tables)
template <typename R>
__global__ void k_init_push(cplx_t<R>* __restrict__ EG, const cplx_t<R>* __restrict__ J, const cplx_t<R>* __restrict__ Gn,
                            const cplx_t<R>* __restrict__ C1, const cplx_t<R>* __restrict__ C2, int64_t n) {
    typedef cplx_t<R> C_;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const C_ a1 = C1[e], a2 = C1[e + n], b1 = C2[e], b2 = C2[e + n];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        const C_ j = J[e + l * n], gn = Gn[e + l * n];
        EG[e + l * n] = cadd(cadd(EG[e + l * n], cmul(a1, j)), cmul(a2, gn));
        EG[e + (3 + l) * n] = cadd(cadd(EG[e + (3 + l) * n], cmul(b1, j)), cmul(b2, gn));
    }
}

// maxwell_solvers.f90:131-164
template <typename R>
__global__ void k_poiss_corr(cplx_t<R>* __restrict__ J, const cplx_t<R>* __restrict__ GD, const cplx_t<R>* __restrict__ Gn,
                             const cplx_t<R>* __restrict__ Gp, double dt_inv, const R* __restrict__ w2, int64_t n) {
    typedef cplx_t<R> C_;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const R w = w2[e];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        const C_ d = crs((R)dt_inv, csub(Gp[e + l * n], Gn[e + l * n]));
        J[e + l * n] = cadd(J[e + l * n], crs(w, cadd(GD[e + l * n], d)));
    }
}

// maxwell_solvers.f90:166-197: j += (grad div j + DT(x) grad rho_n) w2_inv, DT complex per x
template <typename R>
__global__ void k_poiss_corr_stat(cplx_t<R>* __restrict__ J, const cplx_t<R>* __restrict__ GD,
                                  const cplx_t<R>* __restrict__ Gn, const double2* __restrict__ DT,
                                  const R* __restrict__ w2, int nkx, int64_t n) {
    typedef cplx_t<R> C_;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const C_ dt = cconv<R>(DT[e % nkx]);
    const R w = w2[e];
#pragma unroll
    for (int l = 0; l < 3; l++) J[e + l * n] = cadd(J[e + l * n], crs(w, cadd(GD[e + l * n], cmul(dt, Gn[e + l * n]))));
}

// maxwell_solvers.f90:199-226: EG *= exp(fact kx), fact = -0.5 i dt beta0: the
// real part of fact kx is zero, so the factor is (cos, sin) of a*kx with
// a = (-0.5 dt) beta0, formed in FP64
template <typename R>
__global__ void k_field_drift(cplx_t<R>* __restrict__ EG, const double* __restrict__ kx, double a, int nkx, int64_t n) {
    typedef cplx_t<R> C_;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double sn, cs;
    sincos(a * kx[e % nkx], &sn, &cs);
    const C_ p = cmake<R>((R)cs, (R)sn);
#pragma unroll
    for (int l = 0; l < 6; l++) EG[e + l * n] = cmul(EG[e + l * n], p);
}

// maxwell_solvers.f90:228-272
template <typename R>
__global__ void k_mult(cplx_t<R>* __restrict__ v, const R* __restrict__ A, int64_t n, int ncomp) {
    typedef cplx_t<R> C_;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const R a = A[e];
    for (int l = 0; l < ncomp; l++) v[e + l * n] = crs(a, v[e + l * n]);
}

// maxwell_solvers.f90:274-318
template <typename R>
__global__ void k_add(cplx_t<R>* __restrict__ v, const cplx_t<R>* __restrict__ A, int64_t n) {
    typedef cplx_t<R> C_;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    v[e] = cadd(v[e], A[e]);
}

template <typename R>
int32_t psatd_spchrg(cplx_t<R>* EG, const cplx_t<R>* J, const cplx_t<R>* Gn, const cplx_t<R>* Gp, const R* C1,
                     const R* C2, int64_t n) {
    if (n <= 0) return ST_OK;
    KTimer kt(KF_PSATD);
    hipLaunchKernelGGL((k_psatd_spchrg<R>), blocks_for(n, 256), 256, 0, stream(), EG, J, Gn, Gp, C1, C2, n);
    CH_LAUNCH_CHECK("k_psatd_spchrg");
    return ST_OK;
}

template <typename R>
int32_t psatd_wo(cplx_t<R>* EG, const cplx_t<R>* J, const R* C1, const R* C2, int64_t n) {
    if (n <= 0) return ST_OK;
    KTimer kt(KF_PSATD);
    hipLaunchKernelGGL((k_psatd_wo<R>), blocks_for(n, 256), 256, 0, stream(), EG, J, C1, C2, n);
    CH_LAUNCH_CHECK("k_psatd_wo");
    return ST_OK;
}

template <typename R>
int32_t init_push(cplx_t<R>* EG, const cplx_t<R>* J, const cplx_t<R>* Gn, const cplx_t<R>* C1, const cplx_t<R>* C2, int64_t n) {
    if (n <= 0) return ST_OK;
    KTimer kt(KF_PSATD);
    hipLaunchKernelGGL((k_init_push<R>), blocks_for(n, 256), 256, 0, stream(), EG, J, Gn, C1, C2, n);
    CH_LAUNCH_CHECK("k_init_push");
    return ST_OK;
}

template <typename R>
int32_t poiss_corr(cplx_t<R>* J, const cplx_t<R>* GD, const cplx_t<R>* Gn, const cplx_t<R>* Gp, double dt_inv,
                   const R* w2, int64_t n) {
    if (n <= 0) return ST_OK;
    KTimer kt(KF_PSATD);
    hipLaunchKernelGGL((k_poiss_corr<R>), blocks_for(n, 256), 256, 0, stream(), J, GD, Gn, Gp, dt_inv, w2, n);
    CH_LAUNCH_CHECK("k_poiss_corr");
    return ST_OK;
}

template <typename R>
int32_t poiss_corr_stat(cplx_t<R>* J, const cplx_t<R>* GD, const cplx_t<R>* Gn, const double2* DT, const R* w2,
                        int nkx, int64_t n) {
    if (n <= 0) return ST_OK;
    hipLaunchKernelGGL((k_poiss_corr_stat<R>), blocks_for(n, 256), 256, 0, stream(), J, GD, Gn, DT, w2, nkx, n);
    CH_LAUNCH_CHECK("k_poiss_corr_stat");
    return ST_OK;
}

template <typename R>
int32_t field_drift(cplx_t<R>* EG, const double* kx, double beta0, double dt, int nkx, int64_t n) {
    if (n <= 0) return ST_OK;
    hipLaunchKernelGGL((k_field_drift<R>), blocks_for(n, 256), 256, 0, stream(), EG, kx, (-0.5 * dt) * beta0, nkx, n);
    CH_LAUNCH_CHECK("k_field_drift");
    return ST_OK;
}

template <typename R>
int32_t mult_tab(cplx_t<R>* v, const R* A, int64_t n, int ncomp) {
    if (n <= 0) return ST_OK;
    hipLaunchKernelGGL((k_mult<R>), blocks_for(n, 256), 256, 0, stream(), v, A, n, ncomp);
    CH_LAUNCH_CHECK("k_mult");
    return ST_OK;
}

template <typename R>
int32_t add_arr(cplx_t<R>* v, const cplx_t<R>* A, int64_t n) {
    if (n <= 0) return ST_OK;
    hipLaunchKernelGGL((k_add<R>), blocks_for(n, 256), 256, 0, stream(), v, A, n);
    CH_LAUNCH_CHECK("k_add");
    return ST_OK;
}

#define CH_FIELDS_INSTANTIATE(R)                                                                                    \
    template int32_t psatd_spchrg<R>(cplx_t<R>*, const cplx_t<R>*, const cplx_t<R>*, const cplx_t<R>*, const R*,      \
                                     const R*, int64_t);                                                              \
    template int32_t psatd_wo<R>(cplx_t<R>*, const cplx_t<R>*, const R*, const R*, int64_t);                           \
    template int32_t init_push<R>(cplx_t<R>*, const cplx_t<R>*, const cplx_t<R>*, const cplx_t<R>*, const cplx_t<R>*,  \
                                  int64_t);                                                                           \
    template int32_t poiss_corr<R>(cplx_t<R>*, const cplx_t<R>*, const cplx_t<R>*, const cplx_t<R>*, double, const R*, \
                                   int64_t);                                                                          \
    template int32_t mult_tab<R>(cplx_t<R>*, const R*, int64_t, int);                                                  \
    template int32_t add_arr<R>(cplx_t<R>*, const cplx_t<R>*, int64_t);
CH_FIELDS_INSTANTIATE(double)
CH_FIELDS_INSTANTIATE(float)
#undef CH_FIELDS_INSTANTIATE

}  // namespace chimera

using namespace chimera;

static inline int64_t nel(int32_t nkx, int32_t nkr, int32_t nkO) { return (int64_t)nkx * nkr * nkO; }

#define SIZES_OK(nkx, nkr, nkO) \
    if ((nkx) < 0 || (nkr) < 0 || (nkO) < 0) return fail(ST_EINVAL, "negative field dimension")

extern "C" {

int32_t chimera_maxwell_push_with_spchrg(double* EG_fb, const double* j_fb, const double* grad_rho_n_fb,
                                         const double* grad_rho_np1_fb, const double* C1, const double* C2,
                                         int32_t nkx, int32_t nkr, int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    const int64_t n = nel(nkx, nkr, nkO);
    Stage st;
    double2* EG = reinterpret_cast<double2*>(st.inout(EG_fb, 2 * (size_t)n * 6));
    const double2* J = reinterpret_cast<const double2*>(st.in(j_fb, 2 * (size_t)n * 3));
    const double2* Gn = reinterpret_cast<const double2*>(st.in(grad_rho_n_fb, 2 * (size_t)n * 3));
    const double2* Gp = reinterpret_cast<const double2*>(st.in(grad_rho_np1_fb, 2 * (size_t)n * 3));
    const double* a = st.in(C1, (size_t)n * 5);
    const double* b = st.in(C2, (size_t)n * 5);
    CH_TRY(st.status());
    CH_TRY(psatd_spchrg<double>(EG, J, Gn, Gp, a, b, n));
    return st.finish();
}

int32_t chimera_maxwell_push_wo_spchrg(double* EG_fb, const double* j_fb, const double* C1, const double* C2,
                                       int32_t nkx, int32_t nkr, int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    const int64_t n = nel(nkx, nkr, nkO);
    Stage st;
    double2* EG = reinterpret_cast<double2*>(st.inout(EG_fb, 2 * (size_t)n * 6));
    const double2* J = reinterpret_cast<const double2*>(st.in(j_fb, 2 * (size_t)n * 3));
    const double* a = st.in(C1, (size_t)n * 3);
    const double* b = st.in(C2, (size_t)n * 3);
    CH_TRY(st.status());
    CH_TRY(psatd_wo<double>(EG, J, a, b, n));
    return st.finish();
}

int32_t chimera_maxwell_push_wo_spchrg_env(double* EG_fb, const double* j_fb, const double* C1, const double* C2,
                                           int32_t nkx, int32_t nkr, int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    const int64_t n = nel(nkx, nkr, nkO);
    Stage st;
    double2* EG = reinterpret_cast<double2*>(st.inout(EG_fb, 2 * (size_t)n * 6));
    const double2* J = reinterpret_cast<const double2*>(st.in(j_fb, 2 * (size_t)n * 3));
    const double2* a = reinterpret_cast<const double2*>(st.in(C1, 2 * (size_t)n * 3));
    const double2* b = reinterpret_cast<const double2*>(st.in(C2, 2 * (size_t)n * 3));
    CH_TRY(st.status());
    CH_TRY(psatd_wo_cplx<double>(EG, J, a, b, n));
    return st.finish();
}

int32_t chimera_maxwell_init_push(double* EG_fb, const double* j_fb, const double* grad_rho_n_fb, const double* C1,
                                  const double* C2, int32_t nkx, int32_t nkr, int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    const int64_t n = nel(nkx, nkr, nkO);
    Stage st;
    double2* EG = reinterpret_cast<double2*>(st.inout(EG_fb, 2 * (size_t)n * 6));
    const double2* J = reinterpret_cast<const double2*>(st.in(j_fb, 2 * (size_t)n * 3));
    const double2* Gn = reinterpret_cast<const double2*>(st.in(grad_rho_n_fb, 2 * (size_t)n * 3));
    const double2* a = reinterpret_cast<const double2*>(st.in(C1, 2 * (size_t)n * 2));
    const double2* b = reinterpret_cast<const double2*>(st.in(C2, 2 * (size_t)n * 2));
    CH_TRY(st.status());
    CH_TRY(init_push<double>(EG, J, Gn, a, b, n));
    return st.finish();
}

int32_t chimera_poiss_corr(double* j_fb, const double* grad_div_j_fb, const double* grad_rho_n_fb,
                           const double* grad_rho_np1_fb, double dt_inv, const double* w2_inv, int32_t nkx,
                           int32_t nkr, int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    const int64_t n = nel(nkx, nkr, nkO);
    Stage st;
    double2* J = reinterpret_cast<double2*>(st.inout(j_fb, 2 * (size_t)n * 3));
    const double2* GD = reinterpret_cast<const double2*>(st.in(grad_div_j_fb, 2 * (size_t)n * 3));
    const double2* Gn = reinterpret_cast<const double2*>(st.in(grad_rho_n_fb, 2 * (size_t)n * 3));
    const double2* Gp = reinterpret_cast<const double2*>(st.in(grad_rho_np1_fb, 2 * (size_t)n * 3));
    const double* w = st.in(w2_inv, (size_t)n);
    CH_TRY(st.status());
    CH_TRY(poiss_corr<double>(J, GD, Gn, Gp, dt_inv, w, n));
    return st.finish();
}

static int32_t mult_entry(double* v_, const double* A_, int32_t nkx, int32_t nkr, int32_t nkO, int ncomp) {
    SIZES_OK(nkx, nkr, nkO);
    const int64_t n = nel(nkx, nkr, nkO);
    Stage st;
    double2* v = reinterpret_cast<double2*>(st.inout(v_, 2 * (size_t)n * ncomp));
    const double* A = st.in(A_, (size_t)n);
    CH_TRY(st.status());
    CH_TRY(mult_tab<double>(v, A, n, ncomp));
    return st.finish();
}

static int32_t add_entry(double* v_, const double* A_, int32_t nkx, int32_t nkr, int32_t nkO, int ncomp) {
    SIZES_OK(nkx, nkr, nkO);
    const int64_t n = nel(nkx, nkr, nkO) * ncomp;
    Stage st;
    double2* v = reinterpret_cast<double2*>(st.inout(v_, 2 * (size_t)n));
    const double2* A = reinterpret_cast<const double2*>(st.in(A_, 2 * (size_t)n));
    CH_TRY(st.status());
    CH_TRY(add_arr<double>(v, A, n));
    return st.finish();
}

static int32_t dev_only(const char* name, std::initializer_list<const void*> ps);

int32_t chimera_poiss_corr_stat(double* j_fb, const double* grad_div_j_fb, const double* grad_rho_n_fb, const double* DT,
                                const double* w2_inv, int32_t nkx, int32_t nkr, int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    const int64_t n = nel(nkx, nkr, nkO);
    Stage st;
    double2* J = reinterpret_cast<double2*>(st.inout(j_fb, 2 * (size_t)n * 3));
    const double2* GD = reinterpret_cast<const double2*>(st.in(grad_div_j_fb, 2 * (size_t)n * 3));
    const double2* Gn = reinterpret_cast<const double2*>(st.in(grad_rho_n_fb, 2 * (size_t)n * 3));
    const double2* D = reinterpret_cast<const double2*>(st.in(DT, 2 * (size_t)nkx));
    const double* w = st.in(w2_inv, (size_t)n);
    CH_TRY(st.status());
    CH_TRY(poiss_corr_stat<double>(J, GD, Gn, D, w, nkx, n));
    return st.finish();
}

int32_t chimera_field_drift(double* EG_fb, const double* kx, double beta0, double dt, int32_t nkx, int32_t nkr,
                            int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    const int64_t n = nel(nkx, nkr, nkO);
    Stage st;
    double2* EG = reinterpret_cast<double2*>(st.inout(EG_fb, 2 * (size_t)n * 6));
    const double* k = st.in(kx, (size_t)nkx);
    CH_TRY(st.status());
    CH_TRY(field_drift<double>(EG, k, beta0, dt, nkx, n));
    return st.finish();
}

int32_t chimera_poiss_corr_stat_c8(float* j_fb, const float* grad_div_j_fb, const float* grad_rho_n_fb,
                                   const double* DT, const float* w2_inv, int32_t nkx, int32_t nkr, int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    CH_TRY(dev_only("poiss_corr_stat_c8", {j_fb, grad_div_j_fb, grad_rho_n_fb, DT, w2_inv}));
    return poiss_corr_stat<float>(reinterpret_cast<float2*>(j_fb), reinterpret_cast<const float2*>(grad_div_j_fb),
                                  reinterpret_cast<const float2*>(grad_rho_n_fb), reinterpret_cast<const double2*>(DT),
                                  w2_inv, nkx, nel(nkx, nkr, nkO));
}

int32_t chimera_field_drift_c8(float* EG_fb, const double* kx, double beta0, double dt, int32_t nkx, int32_t nkr,
                               int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    CH_TRY(dev_only("field_drift_c8", {EG_fb, kx}));
    return field_drift<float>(reinterpret_cast<float2*>(EG_fb), kx, beta0, dt, nkx, nel(nkx, nkr, nkO));
}

int32_t chimera_omp_mult_vec(double* vec_fb, const double* A, int32_t nkx, int32_t nkr, int32_t nkO) {
    return mult_entry(vec_fb, A, nkx, nkr, nkO, 3);
}
int32_t chimera_omp_mult_scl(double* scl_fb, const double* A, int32_t nkx, int32_t nkr, int32_t nkO) {
    return mult_entry(scl_fb, A, nkx, nkr, nkO, 1);
}
int32_t chimera_omp_add_vec(double* vec_fb, const double* A, int32_t nkx, int32_t nkr, int32_t nkO) {
    return add_entry(vec_fb, A, nkx, nkr, nkO, 3);
}
int32_t chimera_omp_add_scl(double* scl_fb, const double* A, int32_t nkx, int32_t nkr, int32_t nkO) {
    return add_entry(scl_fb, A, nkx, nkr, nkO, 1);
}

/* ---------------------------------------------------------------- resident tier, complex64 (mixed precision) */

static int32_t dev_only(const char* name, std::initializer_list<const void*> ps) {
    for (const void* p : ps)
        if (p && !is_device_ptr(p)) return fail(ST_EINVAL, "%s: resident entry points take device pointers", name);
    return ST_OK;
}

int32_t chimera_maxwell_push_with_spchrg_c8(float* EG_fb, const float* j_fb, const float* grad_rho_n_fb,
                                            const float* grad_rho_np1_fb, const float* C1, const float* C2,
                                            int32_t nkx, int32_t nkr, int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    CH_TRY(dev_only("maxwell_push_with_spchrg_c8", {EG_fb, j_fb, grad_rho_n_fb, grad_rho_np1_fb, C1, C2}));
    return psatd_spchrg<float>(reinterpret_cast<float2*>(EG_fb), reinterpret_cast<const float2*>(j_fb),
                               reinterpret_cast<const float2*>(grad_rho_n_fb),
                               reinterpret_cast<const float2*>(grad_rho_np1_fb), C1, C2, nel(nkx, nkr, nkO));
}

int32_t chimera_maxwell_push_wo_spchrg_c8(float* EG_fb, const float* j_fb, const float* C1, const float* C2,
                                          int32_t nkx, int32_t nkr, int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    CH_TRY(dev_only("maxwell_push_wo_spchrg_c8", {EG_fb, j_fb, C1, C2}));
    return psatd_wo<float>(reinterpret_cast<float2*>(EG_fb), reinterpret_cast<const float2*>(j_fb), C1, C2,
                           nel(nkx, nkr, nkO));
}

int32_t chimera_maxwell_init_push_c8(float* EG_fb, const float* j_fb, const float* grad_rho_n_fb, const float* C1,
                                     const float* C2, int32_t nkx, int32_t nkr, int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    CH_TRY(dev_only("maxwell_init_push_c8", {EG_fb, j_fb, grad_rho_n_fb, C1, C2}));
    return init_push<float>(reinterpret_cast<float2*>(EG_fb), reinterpret_cast<const float2*>(j_fb),
                            reinterpret_cast<const float2*>(grad_rho_n_fb), reinterpret_cast<const float2*>(C1),
                            reinterpret_cast<const float2*>(C2), nel(nkx, nkr, nkO));
}

int32_t chimera_omp_mult_vec_c8(float* vec_fb, const float* A, int32_t nkx, int32_t nkr, int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    CH_TRY(dev_only("omp_mult_vec_c8", {vec_fb, A}));
    return mult_tab<float>(reinterpret_cast<float2*>(vec_fb), A, nel(nkx, nkr, nkO), 3);
}

int32_t chimera_omp_add_vec_c8(float* vec_fb, const float* A, int32_t nkx, int32_t nkr, int32_t nkO) {
    SIZES_OK(nkx, nkr, nkO);
    CH_TRY(dev_only("omp_add_vec_c8", {vec_fb, A}));
    return add_arr<float>(reinterpret_cast<float2*>(vec_fb), reinterpret_cast<const float2*>(A),
                          nel(nkx, nkr, nkO) * 3);
}

}  // extern "C"
