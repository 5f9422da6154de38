"""Drop-in replacement for the reference's f2py module `fimera` on the PIC-step path.

A driver written against the reference (``import chimera.moduls.fimera as chimera``)
switches to the GPU with ``import chimera_amd.fimera as chimera``: every function
below has the f2py signature of the Fortran routine it replaces (cited per
function): positional arguments in Fortran order, intent(hide) dimensions
inferred from array shapes, intent(in,out) arrays updated in place when they are
Fortran-contiguous with the right dtype (else a converted copy is updated) and
returned, several outputs returned as a tuple in declaration order.

The arrays are NumPy host arrays; each call stages them through device memory
and returns when the results are back.  Resident simulations (chimera_amd's
Solver / Specie / ChimeraRun) keep their state on the device and call the same
C ABI with device pointers.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import DEVICE_LOCK, check, lib

__all__ = [
    "push_velocs", "push_coords", "genparts", "sortpartsout", "chunk_coords_boundaries", "align_data_vec",
    "align_data_scl", "dep_curr", "dep_dens", "dep_curr_chnk", "dep_dens_chnk", "proj_fld", "eb_correction",
    "fb_vec_in", "fb_scl_in", "fb_eb_out", "fb_filtr", "fb_rot", "fb_grad", "fb_div", "fb_graddiv",
    "maxwell_push_with_spchrg", "maxwell_push_wo_spchrg", "maxwell_init_push", "poiss_corr", "omp_mult_vec",
    "omp_mult_scl", "omp_add_vec", "omp_add_scl", "undul_analytic", "undul_analytic_taper", "undul_mapped", "undul_mapped_tap", "planewave", "gaussbeam", "poiss_corr_stat", "field_drift",
    "dep_curr_env", "dep_dens_env", "dep_curr_env_chnk", "dep_dens_env_chnk", "proj_fld_env", "eb_correction_env",
    "fb_rot_env", "fb_grad_env", "fb_div_env", "fb_graddiv_env",
]


def _arr(a, dtype):
    a = np.asarray(a)
    if a.dtype == dtype and a.flags.f_contiguous:
        return a
    return np.asfortranarray(a, dtype=dtype)


def _p(a):
    return a.ctypes.data


def _call(name, *args):
    with DEVICE_LOCK:
        st = getattr(lib(), name)(*args)
    check(name, st)


# ------------------------------------------------------------------ particle_tools.f90

def push_velocs(momenta, Fld, dt):
    """particle_tools.f90:18-56 -> momenta"""
    u, F = _arr(momenta, np.float64), _arr(Fld, np.float64)
    _call("chimera_push_velocs", _p(u), _p(F), float(dt), u.shape[1])
    return u


def push_coords(coord, momenta, coord_cntr, dt):
    """particle_tools.f90:58-82 -> (coord, coord_cntr)"""
    x, u, xc = _arr(coord, np.float64), _arr(momenta, np.float64), _arr(coord_cntr, np.float64)
    _call("chimera_push_coords", _p(x), _p(u), _p(xc), float(dt), x.shape[1])
    return x, xc


def genparts(coord, Xgrid, Rgrid, RandPackO, PackX, PackR, PackO):
    """particle_tools.f90:84-128 -> (coord, indPart)"""
    c = _arr(coord, np.float64)
    X, R, RP = _arr(Xgrid, np.float64), _arr(Rgrid, np.float64), _arr(RandPackO, np.float64)
    PX, PR, PO = _arr(PackX, np.float64), _arr(PackR, np.float64), _arr(PackO, np.complex128)
    n = ctypes.c_int32(0)
    _call("chimera_genparts", _p(c), ctypes.addressof(n), _p(X), _p(R), _p(RP), _p(PX), _p(PR), _p(PO), c.shape[1],
          X.shape[0], R.shape[0], PX.shape[0])
    return c, int(n.value)


def sortpartsout(coord, lims):
    """particle_tools.f90:130-153 -> (indx2stay, num2stay)"""
    c, L = _arr(coord, np.float64), _arr(lims, np.float64)
    idx = np.zeros(c.shape[1], dtype=np.int32)
    n = ctypes.c_int32(0)
    _call("chimera_sortpartsout", _p(idx), ctypes.addressof(n), _p(c), _p(L), c.shape[1])
    return idx, int(n.value)


def chunk_coords_boundaries(coord, lims, Xgrid, nchnk):
    """particle_tools.f90:155-208 -> (chunked_indx, IndInChnk, GoOut)"""
    c, L, X = _arr(coord, np.float64), _arr(lims, np.float64), _arr(Xgrid, np.float64)
    ids = np.zeros(c.shape[1], dtype=np.int8)
    ind = np.zeros(int(nchnk) + 1, dtype=np.int32)
    go = ctypes.c_int32(0)
    _call("chimera_chunk_coords_boundaries", _p(ids), _p(ind), ctypes.addressof(go), _p(c), _p(L), _p(X), int(nchnk),
          c.shape[1], X.shape[0] - 1)
    return ids, ind, int(go.value)


def align_data_vec(dat, chunked_indx):
    """particle_tools.f90:270-296 -> dat"""
    d, i = _arr(dat, np.float64), np.ascontiguousarray(chunked_indx, dtype=np.int64)
    _call("chimera_align_data_vec", _p(d), _p(i), i.shape[0], d.shape[1])
    return d


def align_data_scl(dat, chunked_indx):
    """particle_tools.f90:298-324 -> dat"""
    d, i = _arr(dat, np.float64), np.ascontiguousarray(chunked_indx, dtype=np.int64)
    _call("chimera_align_data_scl", _p(d), _p(i), i.shape[0], d.shape[0])
    return d


# ------------------------------------------------------------------ grid_deps.f90

def dep_curr(coord, momenta, wghts, curr, leftX, Rgrid, dx_inv, dr_inv):
    """grid_deps.f90:18-87 -> curr"""
    x, u, w = _arr(coord, np.float64), _arr(momenta, np.float64), _arr(wghts, np.float64)
    J, R = _arr(curr, np.complex128), _arr(Rgrid, np.float64)
    _call("chimera_dep_curr", _p(x), _p(u), _p(w), _p(J), float(leftX), _p(R), float(dx_inv), float(dr_inv),
          x.shape[1], J.shape[0] - 1, J.shape[1] - 1, J.shape[2] - 1)
    return J


def dep_curr_chnk(coord, momenta, wghts, curr, IndInChunk, guards, leftX, Rgrid, dx_inv, dr_inv):
    """grid_deps_chnk.f90:18-130 -> curr"""
    x, u, w = _arr(coord, np.float64), _arr(momenta, np.float64), _arr(wghts, np.float64)
    J, R, I = _arr(curr, np.complex128), _arr(Rgrid, np.float64), _arr(IndInChunk, np.int32)
    _call("chimera_dep_curr_chnk", _p(x), _p(u), _p(w), _p(J), _p(I), int(guards), float(leftX), _p(R),
          float(dx_inv), float(dr_inv), x.shape[1], J.shape[0] - 1, J.shape[1] - 1, J.shape[2] - 1, I.shape[0] - 1)
    return J


def dep_dens(coord, wghts, dens, leftX, Rgrid, dx_inv, dr_inv):
    """grid_deps.f90:89-147 -> dens"""
    x, w = _arr(coord, np.float64), _arr(wghts, np.float64)
    D, R = _arr(dens, np.complex128), _arr(Rgrid, np.float64)
    _call("chimera_dep_dens", _p(x), _p(w), _p(D), float(leftX), _p(R), float(dx_inv), float(dr_inv), x.shape[1],
          D.shape[0] - 1, D.shape[1] - 1, D.shape[2] - 1)
    return D


def dep_dens_chnk(coord, wghts, dens, IndInChunk, guards, leftX, Rgrid, dx_inv, dr_inv):
    """grid_deps_chnk.f90:132-234 -> dens"""
    x, w = _arr(coord, np.float64), _arr(wghts, np.float64)
    D, R, I = _arr(dens, np.complex128), _arr(Rgrid, np.float64), _arr(IndInChunk, np.int32)
    _call("chimera_dep_dens_chnk", _p(x), _p(w), _p(D), _p(I), int(guards), float(leftX), _p(R), float(dx_inv),
          float(dr_inv), x.shape[1], D.shape[0] - 1, D.shape[1] - 1, D.shape[2] - 1, I.shape[0] - 1)
    return D


def proj_fld(coord, wghts, Fld, Fld_tot, leftX, Rgrid, dx_inv, dr_inv):
    """grid_deps.f90:149-217 -> Fld_tot"""
    x, w = _arr(coord, np.float64), _arr(wghts, np.float64)
    F, Ft, R = _arr(Fld, np.complex128), _arr(Fld_tot, np.float64), _arr(Rgrid, np.float64)
    _call("chimera_proj_fld", _p(x), _p(w), _p(F), _p(Ft), float(leftX), _p(R), float(dx_inv), float(dr_inv),
          x.shape[1], F.shape[0] - 1, F.shape[1] - 1, F.shape[2] - 1)
    return Ft


def eb_correction(eb_spc):
    """grid_deps.f90:219-266 -> eb_spc"""
    E = _arr(eb_spc, np.complex128)
    _call("chimera_eb_correction", _p(E), E.shape[0] - 1, E.shape[1] - 1, E.shape[2])
    return E


# ------------------------------------------------------------------ fb_io.f90

def fb_vec_in(vec_fb, vec, leftX, kx, In):
    """fb_io.f90:18-59 -> vec_fb"""
    o, v, k, M = _arr(vec_fb, np.complex128), _arr(vec, np.complex128), _arr(kx, np.float64), _arr(In, np.float64)
    _call("chimera_fb_vec_in", _p(o), _p(v), float(leftX), _p(k), _p(M), v.shape[0], v.shape[1] - 1, v.shape[2],
          M.shape[1])
    return o


def fb_scl_in(scl_fb, scl, leftX, kx, In):
    """fb_io.f90:61-98 -> scl_fb"""
    o, v, k, M = _arr(scl_fb, np.complex128), _arr(scl, np.complex128), _arr(kx, np.float64), _arr(In, np.float64)
    _call("chimera_fb_scl_in", _p(o), _p(v), float(leftX), _p(k), _p(M), v.shape[0], v.shape[1] - 1, v.shape[2],
          M.shape[1])
    return o


def fb_eb_out(eb_spc, e_fb, b_fb, leftX, kx, Out):
    """fb_io.f90:182-228 -> eb_spc"""
    eb, e, b = _arr(eb_spc, np.complex128), _arr(e_fb, np.complex128), _arr(b_fb, np.complex128)
    k, M = _arr(kx, np.float64), _arr(Out, np.float64)
    _call("chimera_fb_eb_out", _p(eb), _p(e), _p(b), float(leftX), _p(k), _p(M), eb.shape[0], eb.shape[1] - 1,
          eb.shape[2], M.shape[0])
    return eb


def fb_filtr(vec, leftX, kx, filtr, modefilt):
    """fb_io.f90:230-308 -> vec"""
    v, k, f = _arr(vec, np.complex128), _arr(kx, np.float64), _arr(filtr, np.float64)
    _call("chimera_fb_filtr", _p(v), float(leftX), _p(k), _p(f), int(modefilt), v.shape[0], v.shape[1], v.shape[2],
          f.shape[0])
    return v


# ------------------------------------------------------------------ fb_math.f90

def _fbop(name, out, inp, DpS2S, DmS2S, kx, out_is_inp=False):
    o = _arr(out, np.complex128)
    i = o if out_is_inp else _arr(inp, np.complex128)
    Dp, Dm, k = _arr(DpS2S, np.float64), _arr(DmS2S, np.float64), _arr(kx, np.float64)
    _call(name, _p(o), *(() if out_is_inp else (_p(i),)), _p(Dp), _p(Dm), _p(k), i.shape[0], i.shape[1],
          i.shape[2] - 1, Dp.shape[1])
    return o


def fb_rot(vec_fb_loc, vec_fb, DpS2S, DmS2S, kx):
    """fb_math.f90:18-94 -> vec_fb_loc"""
    return _fbop("chimera_fb_rot", vec_fb_loc, vec_fb, DpS2S, DmS2S, kx)


def fb_grad(vec_fb_loc, scl_fb, DpS2S, DmS2S, kx):
    """fb_math.f90:96-149 -> vec_fb_loc"""
    return _fbop("chimera_fb_grad", vec_fb_loc, scl_fb, DpS2S, DmS2S, kx)


def fb_div(scl_fb_loc, vec_fb, DpS2S, DmS2S, kx):
    """fb_math.f90:151-199 -> scl_fb_loc"""
    return _fbop("chimera_fb_div", scl_fb_loc, vec_fb, DpS2S, DmS2S, kx)


def fb_graddiv(vec_fb, DpS2S, DmS2S, kx):
    """fb_math.f90:201-293 -> vec_fb"""
    return _fbop("chimera_fb_graddiv", vec_fb, None, DpS2S, DmS2S, kx, out_is_inp=True)


# ------------------------------------------------------------------ grid_deps_env.f90, fb_math_env.f90
# envelope (KxShift) grids: the mode axis runs over -nkO..nkO (length 2 nkO + 1)

def _nko_env(n):
    if n % 2 != 1:
        raise ValueError(f"envelope mode axis of length {n}: the modes are -nkO..nkO (odd length)")
    return (n - 1) // 2


def dep_curr_env(coord, momenta, wghts, curr, leftX, Rgrid, dx_inv, dr_inv, kx0):
    """grid_deps_env.f90:18-95 -> curr"""
    x, u, w = _arr(coord, np.float64), _arr(momenta, np.float64), _arr(wghts, np.float64)
    J, R = _arr(curr, np.complex128), _arr(Rgrid, np.float64)
    _call("chimera_dep_curr_env", _p(x), _p(u), _p(w), _p(J), float(leftX), _p(R), float(dx_inv), float(dr_inv),
          float(kx0), x.shape[1], J.shape[0] - 1, J.shape[1] - 1, _nko_env(J.shape[2]))
    return J


def dep_dens_env(coord, wghts, dens, leftX, Rgrid, dx_inv, dr_inv, kx0):
    """grid_deps_env.f90:97-162 -> dens"""
    x, w = _arr(coord, np.float64), _arr(wghts, np.float64)
    D, R = _arr(dens, np.complex128), _arr(Rgrid, np.float64)
    _call("chimera_dep_dens_env", _p(x), _p(w), _p(D), float(leftX), _p(R), float(dx_inv), float(dr_inv), float(kx0),
          x.shape[1], D.shape[0] - 1, D.shape[1] - 1, _nko_env(D.shape[2]))
    return D


def dep_curr_env_chnk(coord, momenta, wghts, curr, IndInChunk, guards, leftX, Rgrid, dx_inv, dr_inv, kx0):
    """grid_deps_env_chnk.f90:18-145 -> curr"""
    x, u, w = _arr(coord, np.float64), _arr(momenta, np.float64), _arr(wghts, np.float64)
    J, R, I = _arr(curr, np.complex128), _arr(Rgrid, np.float64), _arr(IndInChunk, np.int32)
    _call("chimera_dep_curr_env_chnk", _p(x), _p(u), _p(w), _p(J), _p(I), int(guards), float(leftX), _p(R),
          float(dx_inv), float(dr_inv), float(kx0), x.shape[1], J.shape[0] - 1, J.shape[1] - 1, _nko_env(J.shape[2]),
          I.shape[0] - 1)
    return J


def dep_dens_env_chnk(coord, wghts, dens, IndInChunk, guards, leftX, Rgrid, dx_inv, dr_inv, kx0):
    """grid_deps_env_chnk.f90:147-261 -> dens"""
    x, w = _arr(coord, np.float64), _arr(wghts, np.float64)
    D, R, I = _arr(dens, np.complex128), _arr(Rgrid, np.float64), _arr(IndInChunk, np.int32)
    _call("chimera_dep_dens_env_chnk", _p(x), _p(w), _p(D), _p(I), int(guards), float(leftX), _p(R), float(dx_inv),
          float(dr_inv), float(kx0), x.shape[1], D.shape[0] - 1, D.shape[1] - 1, _nko_env(D.shape[2]),
          I.shape[0] - 1)
    return D


def proj_fld_env(coord, wghts, Fld, Fld_tot, leftX, Rgrid, dx_inv, dr_inv, kx0):
    """grid_deps_env.f90:164-238 -> Fld_tot"""
    x, w = _arr(coord, np.float64), _arr(wghts, np.float64)
    F, Ft, R = _arr(Fld, np.complex128), _arr(Fld_tot, np.float64), _arr(Rgrid, np.float64)
    _call("chimera_proj_fld_env", _p(x), _p(w), _p(F), _p(Ft), float(leftX), _p(R), float(dx_inv), float(dr_inv),
          float(kx0), x.shape[1], F.shape[0] - 1, F.shape[1] - 1, _nko_env(F.shape[2]))
    return Ft


def eb_correction_env(eb_spc):
    """grid_deps_env.f90:240-283 -> eb_spc"""
    E = _arr(eb_spc, np.complex128)
    _call("chimera_eb_correction_env", _p(E), E.shape[0] - 1, E.shape[1] - 1, _nko_env(E.shape[2]))
    return E


def _fbop_env(name, out, inp, DpS2S, DmS2S, kx, out_is_inp=False):
    o = _arr(out, np.complex128)
    i = o if out_is_inp else _arr(inp, np.complex128)
    Dp, Dm, k = _arr(DpS2S, np.float64), _arr(DmS2S, np.float64), _arr(kx, np.float64)
    _call(name, _p(o), *(() if out_is_inp else (_p(i),)), _p(Dp), _p(Dm), _p(k), i.shape[0], i.shape[1],
          _nko_env(i.shape[2]), Dp.shape[1])
    return o


def fb_rot_env(vec_fb_loc, vec_fb, DpS2S, DmS2S, kx):
    """fb_math_env.f90:106-162 -> vec_fb_loc"""
    return _fbop_env("chimera_fb_rot_env", vec_fb_loc, vec_fb, DpS2S, DmS2S, kx)


def fb_grad_env(vec_fb_loc, scl_fb, DpS2S, DmS2S, kx):
    """fb_math_env.f90:18-61 -> vec_fb_loc"""
    return _fbop_env("chimera_fb_grad_env", vec_fb_loc, scl_fb, DpS2S, DmS2S, kx)


def fb_div_env(scl_fb_loc, vec_fb, DpS2S, DmS2S, kx):
    """fb_math_env.f90:63-104 -> scl_fb_loc"""
    return _fbop_env("chimera_fb_div_env", scl_fb_loc, vec_fb, DpS2S, DmS2S, kx)


def fb_graddiv_env(vec_fb, DpS2S, DmS2S, kx):
    """fb_math_env.f90:164-233 -> vec_fb"""
    return _fbop_env("chimera_fb_graddiv_env", vec_fb, None, DpS2S, DmS2S, kx, out_is_inp=True)


# ------------------------------------------------------------------ maxwell_solvers.f90

def maxwell_push_with_spchrg(EG_fb, j_fb, grad_rho_n_fb, grad_rho_np1_fb, C1, C2):
    """maxwell_solvers.f90:18-60 -> EG_fb"""
    EG = _arr(EG_fb, np.complex128)
    j, gn, gp = (_arr(a, np.complex128) for a in (j_fb, grad_rho_n_fb, grad_rho_np1_fb))
    a, b = _arr(C1, np.float64), _arr(C2, np.float64)
    _call("chimera_maxwell_push_with_spchrg", _p(EG), _p(j), _p(gn), _p(gp), _p(a), _p(b), *EG.shape[:3])
    return EG


def maxwell_push_wo_spchrg(EG_fb, j_fb, C1, C2):
    """maxwell_solvers.f90:62-96 -> EG_fb (real PSATD tables of the plain solver, complex ones of the envelope solver)"""
    EG, j = _arr(EG_fb, np.complex128), _arr(j_fb, np.complex128)
    C1, C2 = np.asarray(C1), np.asarray(C2)
    if np.iscomplexobj(C1) or np.iscomplexobj(C2):
        # the envelope solver's complex tables (solvers.py:261-267)
        a, b = _arr(C1, np.complex128), _arr(C2, np.complex128)
        _call("chimera_maxwell_push_wo_spchrg_env", _p(EG), _p(j), _p(a), _p(b), *EG.shape[:3])
        return EG
    a, b = _arr(C1, np.float64), _arr(C2, np.float64)
    _call("chimera_maxwell_push_wo_spchrg", _p(EG), _p(j), _p(a), _p(b), *EG.shape[:3])
    return EG


def maxwell_init_push(EG_fb, j_fb, grad_rho_n_fb, C1, C2):
    """maxwell_solvers.f90:98-129 -> EG_fb"""
    EG, j, gn = _arr(EG_fb, np.complex128), _arr(j_fb, np.complex128), _arr(grad_rho_n_fb, np.complex128)
    a, b = _arr(C1, np.complex128), _arr(C2, np.complex128)
    _call("chimera_maxwell_init_push", _p(EG), _p(j), _p(gn), _p(a), _p(b), *EG.shape[:3])
    return EG


def poiss_corr_stat(j_fb, grad_div_j_fb, grad_rho_n_fb, DT, w2_inv):
    """maxwell_solvers.f90:166-197 -> j_fb"""
    j = _arr(j_fb, np.complex128)
    gd, gn = _arr(grad_div_j_fb, np.complex128), _arr(grad_rho_n_fb, np.complex128)
    dt, w = _arr(DT, np.complex128), _arr(w2_inv, np.float64)
    _call("chimera_poiss_corr_stat", _p(j), _p(gd), _p(gn), _p(dt), _p(w), *j.shape[:3])
    return j


def field_drift(EG_fb, kx, beta0, dt):
    """maxwell_solvers.f90:199-226 -> EG_fb"""
    EG, k = _arr(EG_fb, np.complex128), _arr(kx, np.float64)
    _call("chimera_field_drift", _p(EG), _p(k), float(beta0), float(dt), *EG.shape[:3])
    return EG


def poiss_corr(j_fb, grad_div_j_fb, grad_rho_n_fb, grad_rho_np1_fb, dt_inv, w2_inv):
    """maxwell_solvers.f90:131-164 -> j_fb"""
    j = _arr(j_fb, np.complex128)
    gd, gn, gp = (_arr(a, np.complex128) for a in (grad_div_j_fb, grad_rho_n_fb, grad_rho_np1_fb))
    w = _arr(w2_inv, np.float64)
    _call("chimera_poiss_corr", _p(j), _p(gd), _p(gn), _p(gp), float(dt_inv), _p(w), *j.shape[:3])
    return j


def _elem(name, v, A, adt):
    v, A = _arr(v, np.complex128), _arr(A, adt)
    _call(name, _p(v), _p(A), *v.shape[:3])
    return v


def omp_mult_vec(vec_fb, A):
    """maxwell_solvers.f90:228-250 -> vec_fb"""
    return _elem("chimera_omp_mult_vec", vec_fb, A, np.float64)


def omp_mult_scl(scl_fb, A):
    """maxwell_solvers.f90:252-272 -> scl_fb"""
    return _elem("chimera_omp_mult_scl", scl_fb, A, np.float64)


def omp_add_vec(vec_fb, A):
    """maxwell_solvers.f90:274-296 -> vec_fb"""
    return _elem("chimera_omp_add_vec", vec_fb, A, np.complex128)


def omp_add_scl(scl_fb, A):
    """maxwell_solvers.f90:298-318 -> scl_fb"""
    return _elem("chimera_omp_add_scl", scl_fb, A, np.complex128)


# ------------------------------------------------------------------ devices.f90

def undul_analytic(coord, Fld, t, params):
    """devices.f90:162-203 -> Fld"""
    x, F, p = _arr(coord, np.float64), _arr(Fld, np.float64), _arr(params, np.float64)
    if p.size < 4:
        raise ValueError("undul_analytic: params = (a0, lambda, X0, Lx)")
    _call("chimera_undul_analytic", _p(x), _p(F), float(t), _p(p), x.shape[1])
    return F


def _check_map(name, a0, p, nparams, names):
    if p.size < nparams:
        raise ValueError(f"{name}: params = ({names})")
    if a0.ndim != 2 or a0.shape[0] != 2:
        raise ValueError(f"{name}: a0 must have shape (2, nx), got {a0.shape}")


def undul_mapped(coord, Fld, t, a0, params):
    """devices.f90:18-62 -> Fld"""
    x, F, a, p = (_arr(v, np.float64) for v in (coord, Fld, a0, params))
    _check_map("undul_mapped", a, p, 3, "lambda, Xleft, dx")
    _call("chimera_undul_mapped", _p(x), _p(F), float(t), _p(a), _p(p), x.shape[1], a.shape[1])
    return F


def undul_mapped_tap(coord, Fld, t, a0, params):
    """devices.f90:64-115 -> Fld"""
    x, F, a, p = (_arr(v, np.float64) for v in (coord, Fld, a0, params))
    _check_map("undul_mapped_tap", a, p, 5, "lambda, Xleft, dx, Lx, taper")
    _call("chimera_undul_mapped_tap", _p(x), _p(F), float(t), _p(a), _p(p), x.shape[1], a.shape[1])
    return F


def planewave(coord, Fld, t, params):
    """devices.f90:205-252 -> Fld"""
    x, F, p = _arr(coord, np.float64), _arr(Fld, np.float64), _arr(params, np.float64)
    if p.size < 7:
        raise ValueError("planewave: params = (a0, lambda, X0, Lx, ramp, theta, phi0)")
    _call("chimera_planewave", _p(x), _p(F), float(t), _p(p), x.shape[1])
    return F


def gaussbeam(coord, Fld, time, a0, params):
    """devices.f90:254-297 -> Fld"""
    x, F, p = _arr(coord, np.float64), _arr(Fld, np.float64), _arr(params, np.float64)
    if p.size < 8:
        raise ValueError("gaussbeam: params = (lambda, axis, x0, y0, z0, Lx, Ly, Lz)")
    _call("chimera_gaussbeam", _p(x), _p(F), float(time), float(a0), _p(p), x.shape[1])
    return F


def undul_analytic_taper(coord, Fld, t, params):
    """devices.f90:117-160 -> Fld"""
    x, F, p = _arr(coord, np.float64), _arr(Fld, np.float64), _arr(params, np.float64)
    if p.size < 5:
        raise ValueError("undul_analytic_taper: params = (a0, lambda, X0, Lx, taper)")
    _call("chimera_undul_analytic_taper", _p(x), _p(F), float(t), _p(p), x.shape[1])
    return F
