"""The path bench.py times, against the oracle: RK4 steps of the Cartesian BSSN bundle
replayed from its captured hipGraph (MoL_graph_create / MoL_graph_launch).

One step is four substeps of Ricci_eval_g0/g1 -> the fused RHS + RK4 update kernels
rhs_eval_rk4_substep_g{0,1,2}<S> at S = 0..3 (LDS-staged stencils on 64x1x4 and
16x4x4 tiles) -> the composed one-launch ghost fill -> det(gammabar) enforcement.
The oracle runs the reference's step (MoLtimestepping/C_Code_Generation.py:266-329):
Ricci, RHS, RK4 update loop, the layer-by-layer ghost fill, det(gammabar).

* (14,12,10): even rows (direct global->LDS copies), partial tiles in x and z;
  (13,9,7): odd rows (the copy through registers), partial tiles everywhere;
  (70,9,8): wider than one 64-lane row, so the lane-move kernels (hDD,alpha,cf group)
  run their last row tile moved left over the first, with the in-place last substep.
  6 steps from the seeded flat + Gaussian state.
* 192^3 with seed 1 and dt = 0.5 dx: the benchmark's own state and step, 1 step.

State after the steps: 1e-10 of each gridfunction's interior max norm at interior
points (the evolution tolerance of the other GPU tests), the same at ghost points
after dividing by the composed ghost map's largest coefficient sum (a corner ghost
value is 10 f0 - 15 f1 + 6 f2 in each of three directions, so rounding differences
of the interior reach it multiplied by up to 31^3), and the Hamiltonian-constraint
L2 norm within 1e-10 relative.
"""
import numpy as np
import pytest

from helpers import ghost_gain, interior_ghost_err

pytestmark = pytest.mark.gpu
TOL_EVOL = 1e-10


def _run(torch, Nxx, seed, nsteps):
    from nrpytutorial_amd.runtime import Bundle, Grid, Evolution
    from nrpytutorial_amd.workloads import bssn_gaussian_state_torch
    from oracle_lib import Oracle, OracleGrid
    b = Bundle("bssn_cartesian")
    g = Grid(b, Nxx)
    ev = Evolution(g)
    bssn_gaussian_state_torch(ev.y_n, g.xx_dev, b.evol, seed=seed)
    y = ev.y_n.cpu().numpy()
    dt = 0.5 * g.params.dxx0
    ev.step(dt, nsteps, use_graph=True)
    H = g.interior(ev.hamiltonian())
    torch.cuda.synchronize()
    got = ev.y_n.cpu().numpy()
    l2 = float(torch.sqrt(torch.mean(H * H)))
    ev.free_graph()
    del ev, g
    og = OracleGrid(Oracle("bssn_cartesian"), Nxx, {})
    og.mol_step(y, dt, nsteps)
    H_ref = og.interior(og.hamiltonian(y))
    l2_ref = float(np.sqrt(np.mean(H_ref ** 2)))
    return got, y, l2, l2_ref, ghost_gain(b, Nxx)


@pytest.mark.parametrize("Nxx", [(14, 12, 10), (13, 9, 7), (70, 9, 8)])
def test_graph_rk4_steps_match_oracle(torch_cuda, Nxx):
    print("case: bssn_cartesian MoL_graph 6 RK4 steps vs oracle, Nxx=%s, seed 7" % (Nxx,), flush=True)
    got, ref, l2, l2_ref, gain = _run(torch_cuda, Nxx, 7, 6)
    ei, eg = interior_ghost_err(got, ref, 3, gain, report=True)
    print("interior %.2e, ghosts/%.0f %.2e, H-L2 gpu %.15e oracle %.15e" % (ei, gain, eg, l2, l2_ref))
    assert ei < TOL_EVOL and eg < TOL_EVOL
    assert abs(l2 - l2_ref) <= TOL_EVOL * l2_ref


def test_graph_rk4_step_192_matches_oracle(torch_cuda):
    print("case: bssn_cartesian MoL_graph 1 RK4 step vs oracle, 192^3, seed 1", flush=True)
    got, ref, l2, l2_ref, gain = _run(torch_cuda, (192, 192, 192), 1, 1)
    ei, eg = interior_ghost_err(got, ref, 3, gain, report=True)
    print("192^3 interior %.2e, ghosts/%.0f %.2e, H-L2 gpu %.15e oracle %.15e rel %.2e"
          % (ei, gain, eg, l2, l2_ref, abs(l2 - l2_ref) / l2_ref))
    assert ei < TOL_EVOL and eg < TOL_EVOL
    assert abs(l2 - l2_ref) <= TOL_EVOL * l2_ref
