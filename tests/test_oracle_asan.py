"""Memory safety of the C oracle (host only): its AddressSanitizer + UndefinedBehaviorSanitizer
build (oracle/Makefile target asan, serial) runs its entry points on small grids in a child
Python process with the ASan runtime preloaded: boundary maps and both ghost fills
(extrapolation and parity), initial data, RHSs, RK4 steps, the Sommerfeld condition, det(gammabar)
enforcement and the constraints, in Cartesian, SinhSphericalv2 and Spherical (Brill-Lindquist)
coordinates. Any out-of-bounds access, use after free or undefined behaviour aborts the child.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUNDLES = ["scalarwave_cartesian", "maxwell_cartesian", "scalarwave_sinhsphericalv2", "bssn_spherical_sym2"]

CHILD = r'''
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle_lib import Oracle, OracleGrid

def run(bundle, Nxx, pars, steps, sommerfeld):
    o = Oracle(bundle, variant="asan")
    og = OracleGrid(o, Nxx, pars)
    y = og.zeros(len(o.evol))
    og.initial_data(y)
    og.apply_bcs(y)
    dt = og.find_timestep(0.5)
    og.mol_step(y, dt, steps)
    if sommerfeld:
        og.mol_step_sommerfeld(y, dt, 2)
    assert np.isfinite(y).all(), bundle
    return o, og, y

box = dict(xmin=-3.0, xmax=3.0, ymin=-3.0, ymax=3.0, zmin=-3.0, zmax=3.0)
run("scalarwave_cartesian", (10, 9, 8), dict(box, kk0=0.6, kk1=0.48, kk2=0.64, wavespeed=1.0, time=0.0), 3, True)
run("scalarwave_cartesian", (13, 9, 16), dict(box, kk0=0.6, kk1=0.48, kk2=0.64, wavespeed=1.0, time=0.0), 1, True)
o, og, y = run("maxwell_cartesian", (10, 9, 8), dict(box, time=0.4), 2, True)
og.maxwell_constraints(y)
run("scalarwave_sinhsphericalv2", (12, 8, 10), dict(AMPL=3.0), 2, True)
o, og, y = run("bssn_spherical_sym2", (16, 8, 2),
               dict(BH1_mass=0.5, BH2_mass=0.5, BH1_posn_z=0.5, BH2_posn_z=-0.5, RMAX=7.5, eta=2.0), 2, False)
og.enforce_detg(y)
og.hamiltonian(y)
og.momentum(y)
print("ok")
'''

# Control: the same child with the ghost fill told of one gridfunction more than the array
# holds (a mistake planted in the call, on the host): the sanitizer must report it.
CONTROL = r'''
import sys, ctypes
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle_lib import Oracle, OracleGrid
o = Oracle("scalarwave_cartesian", variant="asan")
og = OracleGrid(o, (10, 9, 8), dict(kk0=0.6, kk1=0.48, kk2=0.64, wavespeed=1.0, time=0.0))
y = og.zeros(1)
o.lib.oracle_apply_bcs(ctypes.byref(og.p), og.bc, 2, o.parity, y.ctypes.data)
print("not caught")
'''


def _libasan():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_under_address_and_ub_sanitizers():
    asan = _libasan()
    if asan is None:
        pytest.skip("gcc has no libasan here")
    gen = os.path.join(ROOT, "oracle", "generated")
    missing = [b for b in BUNDLES if not os.path.exists(os.path.join(gen, b + ".c"))]
    if missing:
        pytest.skip("oracle sources not generated: %s" % missing)
    r = subprocess.run(["make", "-j4", "CC=gcc", "asan", "BUNDLES=" + " ".join(BUNDLES)],
                       cwd=os.path.join(ROOT, "oracle"), capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="1")
    out = subprocess.run([sys.executable, "-c", CHILD, os.path.join(ROOT, "tests")], env=env,
                         capture_output=True, text=True, timeout=600)
    print(out.stdout[-2000:], out.stderr[-4000:])
    assert "AddressSanitizer" not in out.stderr and "runtime error" not in out.stderr, out.stderr[-4000:]
    assert out.returncode == 0, out.stderr[-4000:]
    assert out.stdout.strip().endswith("ok")
    ctl = subprocess.run([sys.executable, "-c", CONTROL, os.path.join(ROOT, "tests")], env=env,
                         capture_output=True, text=True, timeout=600)
    assert "heap-buffer-overflow" in ctl.stderr and ctl.returncode != 0, (ctl.returncode, ctl.stderr[-2000:])
