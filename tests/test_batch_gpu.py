"""Batched Brill-Lindquist evolutions on the GPU: one launch per kernel per
substep for all instances must reproduce, instance by instance, the plain
single-grid evolution (same kernels' arithmetic, same boundary maps). Checked
on a small seeded sweep of uneven grids: final states agree to 1e-13 of each
gridfunction's scale and the Hamiltonian-constraint L2 norms to 1e-12, with
hipGraph replay and after the tile autotuner has changed the tile shape."""
import numpy as np
import pytest

from helpers import rel_maxnorm_err

pytestmark = pytest.mark.gpu


def _sequential(bundle, instances, nsteps):
    from nrpytutorial_amd.runtime import Grid, Evolution
    out = []
    for Nxx, pars in instances:
        g = Grid(bundle, Nxx, **pars)
        ev = Evolution(g)
        ev.initial_data()
        ev.step(g.find_timestep(0.5), nsteps, use_graph=True)
        H = g.interior(ev.hamiltonian())
        out.append((ev.y_n.cpu().numpy(), float((H * H).mean().sqrt())))
        ev.free_graph()
    return out


@pytest.mark.parametrize("tune", [False, True])
def test_batched_matches_sequential(torch_cuda, tune):
    torch = torch_cuda
    from nrpytutorial_amd.runtime import Bundle
    from nrpytutorial_amd.batch import BatchedEvolution, brill_lindquist_sweep, autotune
    print("case: batched BL sweep of 6 instances, 10 steps, autotune=%s" % tune, flush=True)
    b = Bundle("bssn_spherical_sym2")
    inst = brill_lindquist_sweep(6, seed=4, nr=(16, 40), nth=(8, 12))
    nsteps = 10
    be = BatchedEvolution(b, inst)
    if tune:
        best, times = autotune(be, reps=1)
        print("autotuned tile", best, flush=True)
    be.initial_data()
    be.step(nsteps)
    torch.cuda.synchronize()
    got_H = be.hamiltonian_l2()
    got = [ev.y_n.cpu().numpy() for ev in be.evs]
    ref = _sequential(b, inst, nsteps)
    for k, ((y_ref, H_ref), y, H) in enumerate(zip(ref, got, got_H)):
        err = rel_maxnorm_err(y, y_ref, floor_frac=1e-3)
        print("instance %d Nxx=%s: state rel err %.3e, H-L2 %.15e vs %.15e" % (k, inst[k][0], err, H, H_ref))
        assert err < 1e-13
        assert abs(H - H_ref) <= 1e-12 * abs(H_ref)


def test_full_sweep_256_instances_100_steps(torch_cuda):
    """The batched Brill-Lindquist sweep at its stated size: 256 instances (masses in
    [0.3,0.7], separations in [0.6,2.0], Nr in [48,96], Ntheta in [8,16], Nphi = 2,
    RMAX = 7.5), 100 RK4 steps each from one hipGraph of batched launches. Every
    instance's Hamiltonian-constraint L2 norm and final state must be identical (bit for
    bit) to its plain sequential GPU run; four sampled instances (the smallest, the largest
    and two seeded others) must match the oracle's evolution to 1e-10."""
    torch = torch_cuda
    from nrpytutorial_amd.runtime import Bundle
    from nrpytutorial_amd.batch import BatchedEvolution, brill_lindquist_sweep
    from oracle_lib import Oracle, OracleGrid
    print("case: batched BL sweep 256 instances x 100 steps vs sequential and oracle", flush=True)
    b = Bundle("bssn_spherical_sym2")
    inst = brill_lindquist_sweep(256, seed=0)
    sizes = [i[0] for i in inst]
    assert min(s[0] for s in sizes) >= 48 and max(s[0] for s in sizes) <= 96
    assert min(s[1] for s in sizes) >= 8 and max(s[1] for s in sizes) <= 16
    nsteps = 100
    be = BatchedEvolution(b, inst)
    be.initial_data()
    be.step(nsteps)
    torch.cuda.synchronize()
    got_H = be.hamiltonian_l2()
    got = [ev.y_n.cpu().numpy() for ev in be.evs]
    dts = list(be.dts)
    be.free_graph()
    del be
    ref = _sequential(b, inst, nsteps)
    nsame = 0
    for k, ((y_ref, H_ref), y, H) in enumerate(zip(ref, got, got_H)):
        assert H == H_ref, (k, H, H_ref)
        assert np.array_equal(y, y_ref), k
        nsame += 1
    print("%d of 256 instances bit-identical to their sequential runs" % nsame, flush=True)
    npts = [s[0] * s[1] * s[2] for s in sizes]
    picks = sorted({int(np.argmin(npts)), int(np.argmax(npts)), 17, 200})
    o = Oracle("bssn_spherical_sym2")
    for k in picks:
        Nxx, pars = inst[k]
        og = OracleGrid(o, Nxx, pars)
        y = og.zeros(len(o.evol))
        og.initial_data(y)
        og.apply_bcs(y)
        og.enforce_detg(y)
        assert abs(og.find_timestep(0.5) - dts[k]) <= 1e-15 * dts[k]
        og.mol_step(y, dts[k], nsteps)
        err = rel_maxnorm_err(got[k], y, floor_frac=1e-3)
        Ho = og.interior(og.hamiltonian(y))
        l2 = float(np.sqrt(np.mean(Ho ** 2)))
        print("instance %d Nxx=%s: vs oracle state %.2e, H-L2 %.12e vs %.12e" % (k, Nxx, err, got_H[k], l2),
              flush=True)
        assert err < 1e-10
        assert abs(got_H[k] - l2) <= 1e-10 * l2
