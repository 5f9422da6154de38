"""The C ABI: the library loads (no GPU needed to load it) and exports every
entry point include/chimera_hip.h declares, with the argument kinds the header
states; the drop-in module exposes every f2py entry point of the PIC step."""
import ctypes
import os

from chimera_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    decls = _lib.parse_header()
    assert len(decls) >= 50
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [name for name in decls if not hasattr(L, name)]
    assert not missing, missing


def test_header_declares_the_reference_entry_points():
    decls = _lib.parse_header()
    f2py = ["push_velocs", "push_coords", "genparts", "sortpartsout", "chunk_coords_boundaries", "align_data_vec",
            "align_data_scl", "dep_curr", "dep_dens", "dep_curr_chnk", "dep_dens_chnk", "proj_fld", "eb_correction",
            "fb_vec_in", "fb_scl_in", "fb_eb_out", "fb_filtr", "fb_rot", "fb_grad", "fb_div", "fb_graddiv",
            "maxwell_push_with_spchrg", "maxwell_push_wo_spchrg", "maxwell_init_push", "poiss_corr", "omp_mult_vec",
            "omp_mult_scl", "omp_add_vec", "omp_add_scl", "undul_analytic", "dep_curr_env", "dep_dens_env",
            "dep_curr_env_chnk", "dep_dens_env_chnk", "proj_fld_env", "eb_correction_env", "fb_rot_env", "fb_grad_env",
            "fb_div_env", "fb_graddiv_env"]
    for name in f2py:
        assert "chimera_" + name in decls, name
    # scalars by value, arrays by pointer: dep_curr has 8 pointers/doubles then 4 int32 dimensions
    rt, args = decls["chimera_dep_curr"]
    assert args[-4:] == [ctypes.c_int32] * 4
    assert args[4] is ctypes.c_double


def test_dropin_module_mirrors_fimera():
    import chimera_amd.fimera as F
    for name in F.__all__:
        assert callable(getattr(F, name)), name
