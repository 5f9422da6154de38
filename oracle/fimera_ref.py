"""The oracle in the calling convention of the reference's f2py module `fimera`.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg, never by the product (chimera_amd).

Two back ends sit behind the same Python signatures:

* ``Fimera("port")``       -> oracle/_build/libcpu_ref.so, the C restatement in
                              cpu_ref.c (every hot-path routine, FFTs included);
* ``Fimera("reference")``  -> oracle/_ref/libchimera_ref.so, the reference's own
                              FFTW-free Fortran (particle_tools.f90, grid_deps.f90,
                              grid_deps_env.f90, maxwell_solvers.f90, fb_math.f90,
                              fb_math_env.f90, devices.f90)
                              compiled by oracle/Makefile; fb_io.f90 routines are
                              absent from it (FFTW3 is not available).

Calling convention (f2py rules, applied to the `!f2py` directives of each
Fortran routine): positional arguments in Fortran order with intent(hide)
dimensions dropped and inferred from array shapes; intent(in,out) arrays are
updated in place when they are Fortran-contiguous with the right dtype
(otherwise a converted copy is updated) and returned; several outputs come
back as a tuple in declaration order.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PORT_LIB = os.path.join(_HERE, "_build", "libcpu_ref.so")
REF_LIB = os.path.join(_HERE, "_ref", "libchimera_ref.so")
# the port compiled with FMA contraction: the same arithmetic rounded differently
PORT_FMA_LIB = os.path.join(_HERE, "_build", "libcpu_ref_fma.so")

_P = ctypes.c_void_p


def _f(a, dtype):
    """Fortran-contiguous array of dtype; the same object when already so."""
    a = np.asarray(a)
    if a.dtype == dtype and a.flags.f_contiguous:
        return a
    return np.asfortranarray(a, dtype=dtype)


def _ptr(a):
    return a.ctypes.data_as(_P)


class Fimera:
    """f2py-shaped wrapper over one oracle back end ("port", "port_fma" or "reference")."""

    def __init__(self, backend: str = "port"):
        self.backend = backend
        path = {"port": PORT_LIB, "port_fma": PORT_FMA_LIB, "reference": REF_LIB}[backend]
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library {path} is not built (run `make -C oracle` / `make -C oracle ref`)")
        self._lib = ctypes.CDLL(path)
        self._fortran = backend == "reference"

    # -- argument marshalling: the port takes scalars by value, Fortran by reference
    def _call(self, name, *args):
        conv = []
        keep = []
        for a in args:
            if isinstance(a, np.ndarray):
                conv.append(_ptr(a))
            elif isinstance(a, float):
                if self._fortran:
                    v = ctypes.c_double(a)
                    keep.append(v)
                    conv.append(ctypes.byref(v))
                else:
                    conv.append(ctypes.c_double(a))
            elif isinstance(a, (int, np.integer)):
                if self._fortran:
                    v = ctypes.c_int32(int(a))
                    keep.append(v)
                    conv.append(ctypes.byref(v))
                else:
                    conv.append(ctypes.c_int32(int(a)))
            elif isinstance(a, ctypes._SimpleCData):
                keep.append(a)
                conv.append(ctypes.byref(a))
            else:
                raise TypeError(f"unsupported argument {type(a)} for {name}")
        if self._fortran:
            fn = getattr(self._lib, name + "_")
        else:
            fn = getattr(self._lib, "ref_" + name)
        fn.restype = ctypes.c_int
        return fn(*conv)

    def has(self, name: str) -> bool:
        sym = (name + "_") if self._fortran else ("ref_" + name)
        return hasattr(self._lib, sym)

    # ---------------------------------------------------------------- particle_tools.f90
    def push_velocs(self, momenta, Fld, dt):
        momenta = _f(momenta, np.float64)
        Fld = _f(Fld, np.float64)
        self._call("push_velocs", momenta, Fld, float(dt), momenta.shape[1])
        return momenta

    def push_coords(self, coord, momenta, coord_cntr, dt):
        coord, coord_cntr = _f(coord, np.float64), _f(coord_cntr, np.float64)
        momenta = _f(momenta, np.float64)
        self._call("push_coords", coord, momenta, coord_cntr, float(dt), coord.shape[1])
        return coord, coord_cntr

    def genparts(self, coord, Xgrid, Rgrid, RandPackO, PackX, PackR, PackO):
        coord = _f(coord, np.float64)
        Xgrid, Rgrid = _f(Xgrid, np.float64), _f(Rgrid, np.float64)
        RandPackO = _f(RandPackO, np.float64)
        PackX, PackR = _f(PackX, np.float64), _f(PackR, np.float64)
        PackO = _f(PackO, np.complex128)
        ind = ctypes.c_int32(0)
        args = (coord, ind, Xgrid, Rgrid, RandPackO, PackX, PackR, PackO,
                coord.shape[1], Xgrid.shape[0], Rgrid.shape[0], PackX.shape[0])
        if self._fortran:
            self._call("genparts", *args)
        else:
            self._lib.ref_genparts(*[_ptr(a) if isinstance(a, np.ndarray) else
                                     (ctypes.byref(a) if isinstance(a, ctypes.c_int32) else ctypes.c_int32(a))
                                     for a in args])
        return coord, int(ind.value)

    def sortpartsout(self, coord, lims):
        coord, lims = _f(coord, np.float64), _f(lims, np.float64)
        idx = np.zeros(coord.shape[1], dtype=np.int32)
        num = ctypes.c_int32(0)
        if self._fortran:
            self._call("sortpartsout", idx, num, coord, lims, coord.shape[1])
        else:
            self._lib.ref_sortpartsout(_ptr(idx), ctypes.byref(num), _ptr(coord), _ptr(lims),
                                       ctypes.c_int32(coord.shape[1]))
        return idx, int(num.value)

    def chunk_coords_boundaries(self, coord, lims, Xgrid, nchnk):
        coord, lims, Xgrid = _f(coord, np.float64), _f(lims, np.float64), _f(Xgrid, np.float64)
        np_ = coord.shape[1]
        chunked = np.zeros(np_, dtype=np.int8)
        ind = np.zeros(int(nchnk) + 1, dtype=np.int32)
        goout = ctypes.c_int32(0)
        if self._fortran:
            self._call("chunk_coords_boundaries", chunked, ind, goout, coord, lims, Xgrid,
                       int(nchnk), np_, Xgrid.shape[0] - 1)
        else:
            self._lib.ref_chunk_coords_boundaries(_ptr(chunked), _ptr(ind), ctypes.byref(goout), _ptr(coord),
                                                  _ptr(lims), _ptr(Xgrid), ctypes.c_int32(int(nchnk)),
                                                  ctypes.c_int32(np_), ctypes.c_int32(Xgrid.shape[0] - 1))
        return chunked, ind, int(goout.value)

    def _align(self, dat, idx, ncomp):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        if self._fortran:
            name = "align_data_vec" if ncomp == 3 else "align_data_scl"
            self._call(name, dat, idx, idx.shape[0], dat.shape[-1])
        else:
            self._lib.ref_align_data(_ptr(dat), _ptr(idx), ctypes.c_int32(ncomp), ctypes.c_int32(idx.shape[0]),
                                     ctypes.c_int32(dat.shape[-1]))
        return dat

    def align_data_vec(self, dat, chunked_indx):
        return self._align(_f(dat, np.float64), chunked_indx, 3)

    def align_data_scl(self, dat, chunked_indx):
        return self._align(_f(dat, np.float64), chunked_indx, 1)

    # ---------------------------------------------------------------- grid_deps.f90
    def dep_curr(self, coord, momenta, wghts, curr, leftX, Rgrid, dx_inv, dr_inv):
        coord, momenta, wghts = _f(coord, np.float64), _f(momenta, np.float64), _f(wghts, np.float64)
        curr, Rgrid = _f(curr, np.complex128), _f(Rgrid, np.float64)
        nx, nr, nko = curr.shape[0] - 1, curr.shape[1] - 1, curr.shape[2] - 1
        self._call("dep_curr", coord, momenta, wghts, curr, float(leftX), Rgrid, float(dx_inv), float(dr_inv),
                   coord.shape[1], nx, nr, nko)
        return curr

    def dep_dens(self, coord, wghts, dens, leftX, Rgrid, dx_inv, dr_inv):
        coord, wghts = _f(coord, np.float64), _f(wghts, np.float64)
        dens, Rgrid = _f(dens, np.complex128), _f(Rgrid, np.float64)
        nx, nr, nko = dens.shape[0] - 1, dens.shape[1] - 1, dens.shape[2] - 1
        self._call("dep_dens", coord, wghts, dens, float(leftX), Rgrid, float(dx_inv), float(dr_inv),
                   coord.shape[1], nx, nr, nko)
        return dens

    def proj_fld(self, coord, wghts, Fld, Fld_tot, leftX, Rgrid, dx_inv, dr_inv):
        coord, wghts = _f(coord, np.float64), _f(wghts, np.float64)
        Fld, Fld_tot, Rgrid = _f(Fld, np.complex128), _f(Fld_tot, np.float64), _f(Rgrid, np.float64)
        nx, nr, nko = Fld.shape[0] - 1, Fld.shape[1] - 1, Fld.shape[2] - 1
        self._call("proj_fld", coord, wghts, Fld, Fld_tot, float(leftX), Rgrid, float(dx_inv), float(dr_inv),
                   coord.shape[1], nx, nr, nko)
        return Fld_tot

    def eb_correction(self, eb_spc):
        eb = _f(eb_spc, np.complex128)
        self._call("eb_correction", eb, eb.shape[0] - 1, eb.shape[1] - 1, eb.shape[2])
        return eb

    # ---------------------------------------------------------------- fb_io.f90 (port only)
    def fb_vec_in(self, vec_fb, vec, leftX, kx, In):
        vec_fb, vec = _f(vec_fb, np.complex128), _f(vec, np.complex128)
        kx, In = _f(kx, np.float64), _f(In, np.float64)
        self._call("fb_in", vec_fb, vec, float(leftX), kx, In, vec.shape[0], vec.shape[1] - 1, vec.shape[2],
                   In.shape[1], 3)
        return vec_fb

    def fb_scl_in(self, scl_fb, scl, leftX, kx, In):
        scl_fb, scl = _f(scl_fb, np.complex128), _f(scl, np.complex128)
        kx, In = _f(kx, np.float64), _f(In, np.float64)
        self._call("fb_in", scl_fb, scl, float(leftX), kx, In, scl.shape[0], scl.shape[1] - 1, scl.shape[2],
                   In.shape[1], 1)
        return scl_fb

    def fb_eb_out(self, eb_spc, e_fb, b_fb, leftX, kx, Out):
        eb_spc, e_fb, b_fb = _f(eb_spc, np.complex128), _f(e_fb, np.complex128), _f(b_fb, np.complex128)
        kx, Out = _f(kx, np.float64), _f(Out, np.float64)
        self._call("fb_eb_out", eb_spc, e_fb, b_fb, float(leftX), kx, Out, eb_spc.shape[0], eb_spc.shape[1] - 1,
                   eb_spc.shape[2], Out.shape[0])
        return eb_spc

    def fb_filtr(self, vec, leftX, kx, filtr, modefilt):
        vec, kx, filtr = _f(vec, np.complex128), _f(kx, np.float64), _f(filtr, np.float64)
        rc = self._call("fb_filtr", vec, float(leftX), kx, filtr, int(modefilt), vec.shape[0], vec.shape[1],
                        vec.shape[2], filtr.shape[0])
        if rc != 0:
            raise ValueError("fb_filtr: only modefilt=0 ('left') is defined")
        return vec

    # ---------------------------------------------------------------- fb_math.f90
    def fb_rot(self, vec_fb_loc, vec_fb, DpS2S, DmS2S, kx):
        loc, v = _f(vec_fb_loc, np.complex128), _f(vec_fb, np.complex128)
        Dp, Dm, kx = _f(DpS2S, np.float64), _f(DmS2S, np.float64), _f(kx, np.float64)
        self._call("fb_rot", loc, v, Dp, Dm, kx, v.shape[0], v.shape[1], v.shape[2] - 1, loc.shape[1])
        return loc

    def fb_grad(self, vec_fb_loc, scl_fb, DpS2S, DmS2S, kx):
        loc, s = _f(vec_fb_loc, np.complex128), _f(scl_fb, np.complex128)
        Dp, Dm, kx = _f(DpS2S, np.float64), _f(DmS2S, np.float64), _f(kx, np.float64)
        self._call("fb_grad", loc, s, Dp, Dm, kx, s.shape[0], s.shape[1], s.shape[2] - 1, loc.shape[1])
        return loc

    def fb_div(self, scl_fb_loc, vec_fb, DpS2S, DmS2S, kx):
        loc, v = _f(scl_fb_loc, np.complex128), _f(vec_fb, np.complex128)
        Dp, Dm, kx = _f(DpS2S, np.float64), _f(DmS2S, np.float64), _f(kx, np.float64)
        self._call("fb_div", loc, v, Dp, Dm, kx, v.shape[0], v.shape[1], v.shape[2] - 1, loc.shape[1])
        return loc

    def fb_graddiv(self, vec_fb, DpS2S, DmS2S, kx):
        v = _f(vec_fb, np.complex128)
        Dp, Dm, kx = _f(DpS2S, np.float64), _f(DmS2S, np.float64), _f(kx, np.float64)
        self._call("fb_graddiv", v, Dp, Dm, kx, v.shape[0], v.shape[1], v.shape[2] - 1, Dp.shape[1])
        return v

    # ---------------------------------------------------------------- grid_deps_env.f90
    # envelope arrays: mode axis -nkO..nkO (length 2 nkO + 1)
    def dep_curr_env(self, coord, momenta, wghts, curr, leftX, Rgrid, dx_inv, dr_inv, kx0):
        coord, momenta, wghts = _f(coord, np.float64), _f(momenta, np.float64), _f(wghts, np.float64)
        curr, Rgrid = _f(curr, np.complex128), _f(Rgrid, np.float64)
        self._call("dep_curr_env", coord, momenta, wghts, curr, float(leftX), Rgrid, float(dx_inv), float(dr_inv),
                   float(kx0), coord.shape[1], curr.shape[0] - 1, curr.shape[1] - 1, (curr.shape[2] - 1) // 2)
        return curr

    def dep_dens_env(self, coord, wghts, dens, leftX, Rgrid, dx_inv, dr_inv, kx0):
        coord, wghts = _f(coord, np.float64), _f(wghts, np.float64)
        dens, Rgrid = _f(dens, np.complex128), _f(Rgrid, np.float64)
        self._call("dep_dens_env", coord, wghts, dens, float(leftX), Rgrid, float(dx_inv), float(dr_inv),
                   float(kx0), coord.shape[1], dens.shape[0] - 1, dens.shape[1] - 1, (dens.shape[2] - 1) // 2)
        return dens

    def proj_fld_env(self, coord, wghts, Fld, Fld_tot, leftX, Rgrid, dx_inv, dr_inv, kx0):
        coord, wghts = _f(coord, np.float64), _f(wghts, np.float64)
        Fld, Fld_tot, Rgrid = _f(Fld, np.complex128), _f(Fld_tot, np.float64), _f(Rgrid, np.float64)
        self._call("proj_fld_env", coord, wghts, Fld, Fld_tot, float(leftX), Rgrid, float(dx_inv), float(dr_inv),
                   float(kx0), coord.shape[1], Fld.shape[0] - 1, Fld.shape[1] - 1, (Fld.shape[2] - 1) // 2)
        return Fld_tot

    def eb_correction_env(self, eb_spc):
        eb = _f(eb_spc, np.complex128)
        self._call("eb_correction_env", eb, eb.shape[0] - 1, eb.shape[1] - 1, (eb.shape[2] - 1) // 2)
        return eb

    # ---------------------------------------------------------------- fb_math_env.f90
    def _fbop_env(self, name, out, inp, DpS2S, DmS2S, kx):
        out, inp = _f(out, np.complex128), _f(inp, np.complex128)
        Dp, Dm, kx = _f(DpS2S, np.float64), _f(DmS2S, np.float64), _f(kx, np.float64)
        self._call(name, out, inp, Dp, Dm, kx, inp.shape[0], inp.shape[1], (inp.shape[2] - 1) // 2, out.shape[1])
        return out

    def fb_grad_env(self, vec_fb_loc, scl_fb, DpS2S, DmS2S, kx):
        return self._fbop_env("fb_grad_env", vec_fb_loc, scl_fb, DpS2S, DmS2S, kx)

    def fb_div_env(self, scl_fb_loc, vec_fb, DpS2S, DmS2S, kx):
        return self._fbop_env("fb_div_env", scl_fb_loc, vec_fb, DpS2S, DmS2S, kx)

    def fb_rot_env(self, vec_fb_loc, vec_fb, DpS2S, DmS2S, kx):
        return self._fbop_env("fb_rot_env", vec_fb_loc, vec_fb, DpS2S, DmS2S, kx)

    def fb_graddiv_env(self, vec_fb, DpS2S, DmS2S, kx):
        v = _f(vec_fb, np.complex128)
        Dp, Dm, kx = _f(DpS2S, np.float64), _f(DmS2S, np.float64), _f(kx, np.float64)
        self._call("fb_graddiv_env", v, Dp, Dm, kx, v.shape[0], v.shape[1], (v.shape[2] - 1) // 2, Dp.shape[1])
        return v

    # ---------------------------------------------------------------- maxwell_solvers.f90
    def maxwell_push_with_spchrg(self, EG_fb, j_fb, grad_rho_n_fb, grad_rho_np1_fb, C1, C2):
        EG = _f(EG_fb, np.complex128)
        j, gn, gp = _f(j_fb, np.complex128), _f(grad_rho_n_fb, np.complex128), _f(grad_rho_np1_fb, np.complex128)
        C1, C2 = _f(C1, np.float64), _f(C2, np.float64)
        self._call("maxwell_push_with_spchrg", EG, j, gn, gp, C1, C2, EG.shape[0], EG.shape[1], EG.shape[2])
        return EG

    def maxwell_push_wo_spchrg(self, EG_fb, j_fb, C1, C2):
        EG, j = _f(EG_fb, np.complex128), _f(j_fb, np.complex128)
        # the Fortran declares the tables complex (maxwell_solvers.f90:67-68); f2py casts
        # the real tables of the plain solver, the envelope solver's are complex
        cplx = np.iscomplexobj(C1) or np.iscomplexobj(C2)
        dt = np.complex128 if (self._fortran or cplx) else np.float64
        C1, C2 = _f(C1, dt), _f(C2, dt)
        name = "maxwell_push_wo_spchrg_cplx" if (cplx and not self._fortran) else "maxwell_push_wo_spchrg"
        self._call(name, EG, j, C1, C2, EG.shape[0], EG.shape[1], EG.shape[2])
        return EG

    def maxwell_init_push(self, EG_fb, j_fb, grad_rho_n_fb, C1, C2):
        EG, j, gn = _f(EG_fb, np.complex128), _f(j_fb, np.complex128), _f(grad_rho_n_fb, np.complex128)
        C1, C2 = _f(C1, np.complex128), _f(C2, np.complex128)
        self._call("maxwell_init_push", EG, j, gn, C1, C2, EG.shape[0], EG.shape[1], EG.shape[2])
        return EG

    def poiss_corr(self, j_fb, grad_div_j_fb, grad_rho_n_fb, grad_rho_np1_fb, dt_inv, w2_inv):
        j = _f(j_fb, np.complex128)
        gd, gn, gp = (_f(a, np.complex128) for a in (grad_div_j_fb, grad_rho_n_fb, grad_rho_np1_fb))
        w2 = _f(w2_inv, np.float64)
        self._call("poiss_corr", j, gd, gn, gp, float(dt_inv), w2, j.shape[0], j.shape[1], j.shape[2])
        return j

    def poiss_corr_stat(self, j_fb, grad_div_j_fb, grad_rho_n_fb, DT, w2_inv):
        j = _f(j_fb, np.complex128)
        gd, gn, dt = _f(grad_div_j_fb, np.complex128), _f(grad_rho_n_fb, np.complex128), _f(DT, np.complex128)
        w2 = _f(w2_inv, np.float64)
        self._call("poiss_corr_stat", j, gd, gn, dt, w2, j.shape[0], j.shape[1], j.shape[2])
        return j

    def field_drift(self, EG_fb, kx, beta0, dt):
        EG, kx = _f(EG_fb, np.complex128), _f(kx, np.float64)
        self._call("field_drift", EG, kx, float(beta0), float(dt), EG.shape[0], EG.shape[1], EG.shape[2])
        return EG

    def _elem(self, name, v, A, ncomp, adtype):
        v, A = _f(v, np.complex128), _f(A, adtype)
        if self._fortran:
            self._call(name, v, A, v.shape[0], v.shape[1], v.shape[2])
        else:
            fn = "omp_mult" if "mult" in name else "omp_add"
            self._call(fn, v, A, v.shape[0], v.shape[1], v.shape[2], ncomp)
        return v

    def omp_mult_vec(self, vec_fb, A):
        return self._elem("omp_mult_vec", vec_fb, A, 3, np.float64)

    def omp_mult_scl(self, scl_fb, A):
        return self._elem("omp_mult_scl", scl_fb, A, 1, np.float64)

    def omp_add_vec(self, vec_fb, A):
        return self._elem("omp_add_vec", vec_fb, A, 3, np.complex128)

    def omp_add_scl(self, scl_fb, A):
        return self._elem("omp_add_scl", scl_fb, A, 1, np.complex128)

    # ---------------------------------------------------------------- devices.f90
    def undul_analytic(self, coord, Fld, t, params):
        coord, Fld, params = _f(coord, np.float64), _f(Fld, np.float64), _f(params, np.float64)
        self._call("undul_analytic", coord, Fld, float(t), params, coord.shape[1])
        return Fld

    def undul_mapped(self, coord, Fld, t, a0, params):
        coord, Fld, a0, params = (_f(a, np.float64) for a in (coord, Fld, a0, params))
        self._call("undul_mapped", coord, Fld, float(t), a0, params, coord.shape[1], a0.shape[1])
        return Fld

    def undul_mapped_tap(self, coord, Fld, t, a0, params):
        coord, Fld, a0, params = (_f(a, np.float64) for a in (coord, Fld, a0, params))
        self._call("undul_mapped_tap", coord, Fld, float(t), a0, params, coord.shape[1], a0.shape[1])
        return Fld

    def planewave(self, coord, Fld, t, params):
        coord, Fld, params = _f(coord, np.float64), _f(Fld, np.float64), _f(params, np.float64)
        self._call("planewave", coord, Fld, float(t), params, coord.shape[1])
        return Fld

    def gaussbeam(self, coord, Fld, time, a0, params):
        coord, Fld, params = _f(coord, np.float64), _f(Fld, np.float64), _f(params, np.float64)
        self._call("gaussbeam", coord, Fld, float(time), float(a0), params, coord.shape[1])
        return Fld

    def undul_analytic_taper(self, coord, Fld, t, params):
        coord, Fld, params = _f(coord, np.float64), _f(Fld, np.float64), _f(params, np.float64)
        self._call("undul_analytic_taper", coord, Fld, float(t), params, coord.shape[1])
        return Fld
