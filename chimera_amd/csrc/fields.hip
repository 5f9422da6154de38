// fields.hip -- elementwise field kernels of maxwell_solvers.f90: the PSATD
// update with and without space charge, the static initial push, the Poisson
// correction of the current and the multiply/add helpers.  All arrays are
// (nkx, nkr, nkO, ncomp) complex with nkO the number of modes; element e =
// x + nkx*(k + nkr*m) and component stride n = nkx*nkr*nkO.
#include "common.hpp"
#include "../../include/chimera_hip.h"

namespace chimera {

// maxwell_solvers.f90:18-60
__global__ void k_psatd_spchrg(double2* __restrict__ EG, const double2* __restrict__ J,
                               const double2* __restrict__ Gn, const double2* __restrict__ Gp,
                               const double* __restrict__ C1, const double* __restrict__ C2, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double a1 = C1[e], a2 = C1[e + n], a3 = C1[e + 2 * n], a4 = C1[e + 3 * n], a5 = C1[e + 4 * n];
    const double b1 = C2[e], b2 = C2[e + n], b3 = C2[e + 2 * n], b4 = C2[e + 3 * n], b5 = C2[e + 4 * n];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        const double2 E = EG[e + l * n], G = EG[e + (3 + l) * n];
        const double2 j = J[e + l * n], gn = Gn[e + l * n], gp = Gp[e + l * n];
        EG[e + l * n] = cadd(cadd(cadd(cadd(crs(a1, E), crs(a2, G)), crs(a3, j)), crs(a4, gn)), crs(a5, gp));
        EG[e + (3 + l) * n] = cadd(cadd(cadd(cadd(crs(b1, E), crs(b2, G)), crs(b3, j)), crs(b4, gn)), crs(b5, gp));
    }
}

// maxwell_solvers.f90:62-96 (real tables)
__global__ void k_psatd_wo(double2* __restrict__ EG, const double2* __restrict__ J, const double* __restrict__ C1,
                           const double* __restrict__ C2, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double a1 = C1[e], a2 = C1[e + n], a3 = C1[e + 2 * n];
    const double b1 = C2[e], b2 = C2[e + n], b3 = C2[e + 2 * n];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        const double2 E = EG[e + l * n], G = EG[e + (3 + l) * n], j = J[e + l * n];
        EG[e + l * n] = cadd(cadd(crs(a1, E), crs(a2, G)), crs(a3, j));
        EG[e + (3 + l) * n] = cadd(cadd(crs(b1, E), crs(b2, G)), crs(b3, j));
    }
}

// maxwell_solvers.f90:98-129 (complex tables)
__global__ void k_init_push(double2* __restrict__ EG, const double2* __restrict__ J, const double2* __restrict__ Gn,
                            const double2* __restrict__ C1, const double2* __restrict__ C2, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double2 a1 = C1[e], a2 = C1[e + n], b1 = C2[e], b2 = C2[e + n];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        const double2 j = J[e + l * n], gn = Gn[e + l * n];
        EG[e + l * n] = cadd(cadd(EG[e + l * n], cmul(a1, j)), cmul(a2, gn));
        EG[e + (3 + l) * n] = cadd(cadd(EG[e + (3 + l) * n], cmul(b1, j)), cmul(b2, gn));
    }
}

// maxwell_solvers.f90:131-164
__global__ void k_poiss_corr(double2* __restrict__ J, const double2* __restrict__ GD, const double2* __restrict__ Gn,
                             const double2* __restrict__ Gp, double dt_inv, const double* __restrict__ w2, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double w = w2[e];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        const double2 d = crs(dt_inv, csub(Gp[e + l * n], Gn[e + l * n]));
        J[e + l * n] = cadd(J[e + l * n], crs(w, cadd(GD[e + l * n], d)));
    }
}

// maxwell_solvers.f90:228-272
__global__ void k_mult(double2* __restrict__ v, const double* __restrict__ A, int64_t n, int ncomp) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double a = A[e];
    for (int l = 0; l < ncomp; l++) v[e + l * n] = crs(a, v[e + l * n]);
}

// maxwell_solvers.f90:274-318
__global__ void k_add(double2* __restrict__ v, const double2* __restrict__ A, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    v[e] = cadd(v[e], A[e]);
}

int32_t psatd_spchrg(double2* EG, const double2* J, const double2* Gn, const double2* Gp, const double* C1,
                     const double* C2, int64_t n) {
    if (n <= 0) return ST_OK;
    KTimer kt(KF_PSATD);
    hipLaunchKernelGGL(k_psatd_spchrg, blocks_for(n, 256), 256, 0, stream(), EG, J, Gn, Gp, C1, C2, n);
    CH_LAUNCH_CHECK("k_psatd_spchrg");
    return ST_OK;
}

int32_t psatd_wo(double2* EG, const double2* J, const double* C1, const double* C2, int64_t n) {
    if (n <= 0) return ST_OK;
    KTimer kt(KF_PSATD);
    hipLaunchKernelGGL(k_psatd_wo, blocks_for(n, 256), 256, 0, stream(), EG, J, C1, C2, n);
    CH_LAUNCH_CHECK("k_psatd_wo");
    return ST_OK;
}

int32_t init_push(double2* EG, const double2* J, const double2* Gn, const double2* C1, const double2* C2, int64_t n) {
    if (n <= 0) return ST_OK;
    KTimer kt(KF_PSATD);
    hipLaunchKernelGGL(k_init_push, blocks_for(n, 256), 256, 0, stream(), EG, J, Gn, C1, C2, n);
    CH_LAUNCH_CHECK("k_init_push");
    return ST_OK;
}

int32_t poiss_corr(double2* J, const double2* GD, const double2* Gn, const double2* Gp, double dt_inv,
                   const double* w2, int64_t n) {
    if (n <= 0) return ST_OK;
    KTimer kt(KF_PSATD);
    hipLaunchKernelGGL(k_poiss_corr, blocks_for(n, 256), 256, 0, stream(), J, GD, Gn, Gp, dt_inv, w2, n);
    CH_LAUNCH_CHECK("k_poiss_corr");
    return ST_OK;
}

int32_t mult_tab(double2* v, const double* A, int64_t n, int ncomp) {
    if (n <= 0) return ST_OK;
    hipLaunchKernelGGL(k_mult, blocks_for(n, 256), 256, 0, stream(), v, A, n, ncomp);
    CH_LAUNCH_CHECK("k_mult");
    return ST_OK;
}

int32_t add_arr(double2* v, const double2* A, int64_t n) {
    if (n <= 0) return ST_OK;
    hipLaunchKernelGGL(k_add, blocks_for(n, 256), 256, 0, stream(), v, A, n);
    CH_LAUNCH_CHECK("k_add");
    return ST_OK;
}

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
    CH_TRY(psatd_spchrg(EG, J, Gn, Gp, a, b, n));
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
    CH_TRY(psatd_wo(EG, J, a, b, n));
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
    CH_TRY(init_push(EG, J, Gn, a, b, n));
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
    CH_TRY(poiss_corr(J, GD, Gn, Gp, dt_inv, w, n));
    return st.finish();
}

static int32_t mult_entry(double* v_, const double* A_, int32_t nkx, int32_t nkr, int32_t nkO, int ncomp) {
    SIZES_OK(nkx, nkr, nkO);
    const int64_t n = nel(nkx, nkr, nkO);
    Stage st;
    double2* v = reinterpret_cast<double2*>(st.inout(v_, 2 * (size_t)n * ncomp));
    const double* A = st.in(A_, (size_t)n);
    CH_TRY(st.status());
    CH_TRY(mult_tab(v, A, n, ncomp));
    return st.finish();
}

static int32_t add_entry(double* v_, const double* A_, int32_t nkx, int32_t nkr, int32_t nkO, int ncomp) {
    SIZES_OK(nkx, nkr, nkO);
    const int64_t n = nel(nkx, nkr, nkO) * ncomp;
    Stage st;
    double2* v = reinterpret_cast<double2*>(st.inout(v_, 2 * (size_t)n));
    const double2* A = reinterpret_cast<const double2*>(st.in(A_, 2 * (size_t)n));
    CH_TRY(st.status());
    CH_TRY(add_arr(v, A, n));
    return st.finish();
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

}  // extern "C"
