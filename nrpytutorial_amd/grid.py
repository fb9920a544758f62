"""Gridfunction registry and memory layout.

Follows grid.py (grid.py:12-268):

* gridfunctions are registered as EVOL, AUX or AUXEVOL (grid.py:145-193);
* the memory index of a gridfunction is its position in the
  case-sensitive alphabetical order of its group (grid.py:226-243), and the
  layout is structure-of-arrays with the gridfunction index slowest:
  ``IDX4S(g,i,j,k) = i + N0*(j + N1*(k + N2*g))``;
* ``gfaccess`` spells an access to a gridfunction (grid.py:62-111).

Nxx, Nxx_plus_2NGHOSTS, xx, dxx and invdx are runtime C parameters, as at
grid.py:21-29.
"""
from collections import namedtuple
import sympy as sp
from . import NRPy_param_funcs as par

thismodule = "grid"
par.initialize_param(par.glb_param("char", thismodule, "GridFuncMemAccess", "SENRlike"))
par.initialize_param(par.glb_param("char", thismodule, "MemAllocStyle", "210"))
par.initialize_param(par.glb_param("int", thismodule, "DIM", 3))

glb_gridfc = namedtuple("gridfunction", "gftype name rank DIM")
glb_gridfcs_list = []

Nxx = par.Cparameters("int", thismodule, ["Nxx0", "Nxx1", "Nxx2"], [64, 32, 64])
Nxx_plus_2NGHOSTS = par.Cparameters("int", thismodule,
                                    ["Nxx_plus_2NGHOSTS0", "Nxx_plus_2NGHOSTS1", "Nxx_plus_2NGHOSTS2"],
                                    [70, 38, 70])
xx = par.Cparameters("REAL", thismodule, ["xx0", "xx1", "xx2"], 1e300)
dxx = par.Cparameters("REAL", thismodule, ["dxx0", "dxx1", "dxx2"], 0.1)
invdx = par.Cparameters("REAL", thismodule, ["invdx0", "invdx1", "invdx2"], 1.0)

GF_TYPES = ("EVOL", "AUX", "AUXEVOL")


def reset():
    """Forget every registered gridfunction (tests call this between modules)."""
    del glb_gridfcs_list[:]


def variable_type(var):
    """'gridfunction', 'Cparameter' or 'other' (grid.py:31-53)."""
    name = str(var)
    is_gf = any(gf.name == name for gf in glb_gridfcs_list)
    is_par = par.is_Cparameter(name)
    if is_gf and is_par:
        raise ValueError("variable " + name + " is registered both as a gridfunction and as a Cparameter")
    if is_gf:
        return "gridfunction"
    if is_par:
        return "Cparameter"
    return "other"


def find_gftype(varname):
    for gf in glb_gridfcs_list:
        if gf.name == varname:
            return gf.gftype
    raise KeyError("gridfunction " + varname + " not registered")


def gfaccess(gfarrayname="", varname="", ijklstring=""):
    """Access string for a gridfunction, by grid::GridFuncMemAccess (grid.py:62-111).

    * ``SENRlike``: one array per group, ``in_gfs[IDX4(HDD00GF, i0,i1,i2)]``; AUX and
      AUXEVOL gridfunctions always live in ``aux_gfs`` / ``auxevol_gfs`` (grid.py:80-85).
    * ``ETK``: one array per gridfunction, named as the Einstein Toolkit declares it,
      ``hDD00GF[CCTK_GFINDEX3D(cctkGH, i0,i1,i2)]``, and ``hDD00_rhsGF[...]`` when the
      array is ``rhs_gfs`` (grid.py:92-100). Three dimensions only."""
    gftype = find_gftype(varname)
    idx = ijklstring if ijklstring else "i0, i1, i2"
    mode = par.parval_from_str("GridFuncMemAccess")
    if mode == "ETK":
        if par.parval_from_str("DIM") != 3:
            raise ValueError("GridFuncMemAccess = ETK requires three-dimensional gridfunctions")
        return varname + ("_rhsGF" if gfarrayname == "rhs_gfs" else "GF") + "[CCTK_GFINDEX3D(cctkGH, " + idx + ")]"
    if mode != "SENRlike":
        raise ValueError("grid::GridFuncMemAccess = " + mode + " not supported")
    if gftype == "AUX":
        gfarrayname = "aux_gfs"
    elif gftype == "AUXEVOL":
        gfarrayname = "auxevol_gfs"
    if not gfarrayname:
        raise ValueError("gfaccess needs an array name")
    return gfarrayname + "[IDX4(" + varname.upper() + "GF, " + idx + ")]"


def verify_gridfunction_basename_is_valid(gf_basename):
    if not isinstance(gf_basename, str) or not gf_basename:
        raise ValueError("gridfunction names must be non-empty strings")
    if gf_basename[-1].isdigit():
        raise ValueError("gridfunction base name may not end in a digit: " + gf_basename)


def register_gridfunctions(gf_type, gf_names, rank=0, is_indexed=False, DIM=3):
    """Register gridfunctions and return their SymPy symbols (grid.py:145-193)."""
    if (rank > 0) != bool(is_indexed):
        raise ValueError("indexed gridfunctions must have rank>0 and scalars rank 0")
    if not isinstance(gf_names, list):
        gf_names = [gf_names]
    if not is_indexed:
        for n in gf_names:
            verify_gridfunction_basename_is_valid(n)
    if gf_type not in GF_TYPES:
        raise ValueError("unsupported gridfunction type " + gf_type)
    for n in gf_names:
        if any(gf.name == n for gf in glb_gridfcs_list):
            raise ValueError("gridfunction " + n + " registered twice")
        glb_gridfcs_list.append(glb_gridfc(gf_type, n, rank, DIM))
    syms = [sp.Symbol(n, real=True) for n in gf_names]
    return syms[0] if len(syms) == 1 else syms


def gridfunction_lists():
    """Alphabetically sorted EVOL, AUX, AUXEVOL name lists (grid.py:195-212)."""
    out = []
    for t in GF_TYPES:
        out.append(sorted(gf.name for gf in glb_gridfcs_list if gf.gftype == t))
    return out[0], out[1], out[2]


def gf_index(name):
    """Memory index of a gridfunction inside its group's array."""
    gftype = find_gftype(name)
    evol, aux, auxevol = gridfunction_lists()
    return {"EVOL": evol, "AUX": aux, "AUXEVOL": auxevol}[gftype].index(name)


def gridfunction_defines():
    """Text of gridfunction_defines.h (grid.py:245-268)."""
    evol, aux, auxevol = gridfunction_lists()
    s = "/* Gridfunction indices: alphabetical within each group. */\n"
    for label, lst in (("EVOL", evol), ("AUX", aux), ("AUXEVOL", auxevol)):
        s += "#define NUM_%s_GFS %d\n" % (label, len(lst))
        for i, n in enumerate(lst):
            s += "#define %sGF\t%d\n" % (n.upper(), i)
    return s


def output__gridfunction_defines_h__return_gf_lists(outdir):
    import os
    with open(os.path.join(outdir, "gridfunction_defines.h"), "w") as f:
        f.write(gridfunction_defines())
    return gridfunction_lists()
