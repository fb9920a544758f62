"""Start-to-Finish driver: the time loop and output files of the reference's
Brill-Lindquist head-on collision main()
(Tutorial-Start_to_Finish-BSSNCurvilinear-Two_BHs_Collide.ipynb:975-1187).

* dt = find_timestep(CFL_FACTOR) over interior points, N_final =
  (int)(t_final/dt + 0.5), and the loop runs n = 0 .. N_final inclusive;
* every ``output_every`` iterations, before the step, ``out<Nxx0>-<n:08d>.txt``;
  after the step of iteration N_final - 1, ``out<Nxx0>.txt``;
* each file has one line per interior point, i2 slowest then i1 then i0, with
  ``%e %e %e %e`` of Cartesian y, Cartesian z, the conformal factor cf and
  log10|H| (Hamiltonian constraint), exactly as the reference's fprintf.

The state lives on the GPU; the two output gridfunctions are copied to the host
for writing.
"""
import ctypes
import math
import os
import time

import numpy as np

from .runtime import Bundle, Grid, Evolution, NrpyError


def format_out2D(ycart, zcart, cf, H):
    """The reference's output lines for flattened per-point arrays (C printf %e)."""
    lines = []
    for y, z, c, h in zip(ycart, zcart, cf, H):
        a = abs(h)
        lh = math.log10(a) if a > 0 else -math.inf
        lines.append("%e %e %e %e\n" % (y, z, c, lh))
    return "".join(lines)


def cartesian_yz(grid):
    """Cartesian y and z of every interior point, in output order (i2, i1, i0)."""
    b = grid.bundle
    G = b.NGHOSTS
    xx = grid.xx
    ptrs = (ctypes.c_void_p * 3)(*[x.ctypes.data for x in xx])
    N = [len(x) for x in xx]
    out = np.zeros(3)
    ys, zs = [], []
    for i2 in range(G, N[2] - G):
        for i1 in range(G, N[1] - G):
            for i0 in range(G, N[0] - G):
                b.lib.xx_to_Cart(ctypes.byref(grid.params), ptrs, i0, i1, i2, out.ctypes.data_as(
                    ctypes.POINTER(ctypes.c_double)))
                ys.append(out[1])
                zs.append(out[2])
    return np.array(ys), np.array(zs)


def write_out2D(ev, path, yz=None):
    """Write the current state of ``ev`` (a BSSN Evolution) in the reference's format."""
    g = ev.g
    H = g.interior(ev.hamiltonian()).cpu().numpy().ravel()
    cf = g.interior(ev.y_n[g.bundle.gf("cf")]).cpu().numpy().ravel()
    ycart, zcart = yz if yz is not None else cartesian_yz(g)
    with open(path, "w") as f:
        f.write(format_out2D(ycart, zcart, cf, H))


def two_bhs_collide(Nxx=(72, 12, 2), CFL_FACTOR=0.5, t_final=None, outdir=".", output_every=100,
                    bundle="bssn_spherical_sym2", **params):
    """Run the head-on Brill-Lindquist collision; returns a summary dict. Defaults follow the
    reference's notebook (m1 = m2 = 0.5 at z = +-0.5, RMAX = 7.5 = t_final, CFL 0.5)."""
    pars = dict(BH1_mass=0.5, BH2_mass=0.5, BH1_posn_z=0.5, BH2_posn_z=-0.5, RMAX=7.5, eta=2.0)
    pars.update(params)
    b = Bundle(bundle)
    if b.meta.get("family", "bssn") != "bssn" and not bundle.startswith("bssn"):
        raise NrpyError("two_bhs_collide needs a BSSN bundle")
    g = Grid(b, Nxx, **pars)
    ev = Evolution(g)
    ev.initial_data()
    t_final = pars["RMAX"] if t_final is None else t_final
    dt = g.find_timestep(CFL_FACTOR)
    N_final = int(t_final / dt + 0.5)
    yz = cartesian_yz(g)
    os.makedirs(outdir, exist_ok=True)
    files = []
    t0 = time.perf_counter()
    for n in range(N_final + 1):
        if n % output_every == 0:
            p = os.path.join(outdir, "out%d-%08d.txt" % (Nxx[0], n))
            write_out2D(ev, p, yz)
            files.append(p)
        ev.step(dt, 1)
        if n == N_final - 1:
            p = os.path.join(outdir, "out%d.txt" % Nxx[0])
            write_out2D(ev, p, yz)
            files.append(p)
    g.torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    npts = Nxx[0] * Nxx[1] * Nxx[2]
    return {"dt": dt, "N_final": N_final, "files": files, "seconds": seconds,
            "gp_per_s": 4.0 * npts * (N_final + 1) / seconds, "evolution": ev}
