"""MoL_C_Code_Generation (MoLtimestepping/C_Code_Generation.py:53-335): the files it
writes, the placeholder substitution of the diagonal RK4 schedule, and its refusal of
methods it does not implement. The generated step itself runs in every bundle built
with the Sommerfeld condition (MoL_step_forward_in_time_sommerfeld, GPU tests)."""
import os

import pytest

from nrpytutorial_amd.MoLtimestepping import C_Code_Generation as MoL

RHS = ("rhs_eval(params, RK_INPUT_GFS, RK_OUTPUT_GFS, stream);\n"
       " apply_bcs_sommerfeld(params, xx, bcstruct, NUM_EVOL_GFS, evol_gf_parity, RK_INPUT_GFS, RK_OUTPUT_GFS, stream);")
POST = "apply_bcs_curvilinear(params, bcstruct, NUM_EVOL_GFS, evol_gf_parity, RK_OUTPUT_GFS, stream);"


def test_files_and_schedule(tmp_path):
    MoL.MoL_C_Code_Generation("RK4", RHS_string=RHS, post_RHS_string=POST, outdir=str(tmp_path))
    for f in ("RK_Allocate_Memory.h", "RK_Free_Memory.h", "RK_MoL.h"):
        assert os.path.exists(tmp_path / f)
    mol = (tmp_path / "RK_MoL.h").read_text()
    assert "RK_INPUT_GFS" not in mol and "RK_OUTPUT_GFS" not in mol
    # diagonal schedule: RHS(in) -> out, update, post-RHS on the new stage (y_n after the last)
    want = [("y_n_gfs", "k_odd_gfs", "k_odd_gfs"), ("k_odd_gfs", "k_even_gfs", "k_even_gfs"),
            ("k_even_gfs", "k_odd_gfs", "k_odd_gfs"), ("k_odd_gfs", "k_even_gfs", "y_n_gfs")]
    pos = 0
    for s, (i, o, post) in enumerate(want):
        for needle in ("rhs_eval(params, %s, %s, stream)" % (i, o),
                       "apply_bcs_sommerfeld(params, xx, bcstruct, NUM_EVOL_GFS, evol_gf_parity, %s, %s, stream)" % (i, o),
                       "nrpy_rk4_update_all(%d, (size_t)NUM_EVOL_GFS * Ntot, %s, y_n_gfs, y_nplus1_running_total_gfs, dt, "
                       "stream)" % (s, o),
                       "apply_bcs_curvilinear(params, bcstruct, NUM_EVOL_GFS, evol_gf_parity, %s, stream)" % post):
            k = mol.find(needle, pos)
            assert k >= 0, needle
            pos = k
    assert mol.count("if ((rc = ") == 16
    alloc = (tmp_path / "RK_Allocate_Memory.h").read_text()
    for a in ("y_n_gfs", "k_odd_gfs", "k_even_gfs", "y_nplus1_running_total_gfs"):
        assert "hipMalloc((void **)&%s" % a in alloc


def test_mem_alloc_only(tmp_path):
    MoL.MoL_C_Code_Generation("RK4", outdir=str(tmp_path), MemAllocOnly=True)
    assert os.path.exists(tmp_path / "RK_Allocate_Memory.h") and not os.path.exists(tmp_path / "RK_MoL.h")


def test_other_methods_refused(tmp_path):
    with pytest.raises(ValueError):
        MoL.MoL_C_Code_Generation("RK3 Heun", RHS_string=RHS, outdir=str(tmp_path))
