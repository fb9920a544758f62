"""Generate the system bundles, calling the codegen API the way the reference's
Start-to-Finish drivers do.

* ``scalarwave_cartesian``: Tutorial-Start_to_Finish-ScalarWave.ipynb
  (rhs_eval from ScalarWave_RHSs, plane-wave initial data, quadratic
  extrapolation outer boundaries).
* ``bssn_<coord>``: Tutorial-Start_to_Finish-BSSNCurvilinear-Two_BHs_Collide.ipynb:
  BSSN_RHSs + gauge RHSs with LeaveRicciSymbolic and upwinding on beta^i
  (:436-484), Ricci_eval (:486-540), Hamiltonian_constraint (:542-556),
  enforce_detgammabar_constraint (Enforce_Detgammabar_Constraint.py:60-75),
  Brill-Lindquist initial_data (BSSN_ID_function_string.py:9-39), rfm
  precompute on, symmetry axis "2" for the axisymmetric Spherical runs.
* ``scalarwave_sinhsphericalv2`` / ``scalarwave_sinhcylindricalv2``:
  Tutorial-Start_to_Finish-ScalarWaveCurvilinear.ipynb (ScalarWaveCurvilinear_RHSs
  with rfm precompute, plane-wave data) in the two coordinate systems of the
  reference's curvilinear boundary-condition validation
  (Tutorial-Start_to_Finish-Curvilinear_BCs.ipynb).
* ``maxwell_cartesian``: VacuumMaxwell_Flat_Evol_Cartesian System II.

Each bundle is generated in a fresh interpreter (the registries are global):
``python -m nrpytutorial_amd.codegen.systems <bundle>``.
"""
import re
import sys
import time
import sympy as sp

from .. import NRPy_param_funcs as par
from .. import grid as gri
from .. import indexedexp as ixp
from .. import reference_metric as rfm
from .. import finite_difference as fin
from .. import outputC as oc
from ..outputC import lhrh
from . import lds
from . import addressing
from . import bundle as bd
from . import adjoint as adj
from . import batch as bt

CURVI_SCALARWAVE = {"scalarwave_sinhsphericalv2": "SinhSphericalv2",
                    "scalarwave_sinhcylindricalv2": "SinhCylindricalv2",
                    "scalarwave_symtp": "SymTP",
                    "scalarwave_sinhspherical": "SinhSpherical",
                    "scalarwave_sinhcylindrical": "SinhCylindrical",
                    "scalarwave_sinhsymtp": "SinhSymTP"}
CURVI_MAXWELL = {"maxwell_sinhcylindrical": "SinhCylindrical"}
BUNDLES = ("scalarwave_cartesian", "bssn_cartesian", "bssn_spherical_sym2", "maxwell_cartesian") + \
    tuple(CURVI_SCALARWAVE) + ("bssn_etk",) + tuple(CURVI_MAXWELL)


def _mol_glue(has_ricci, has_detg, rhs_args_decl, rhs_call_args):
    """MoL_step_forward_in_time + hipGraph replay for the diagonal RK4 schedule."""
    ricci = ("  if ((rc = Ricci_eval(rfmstruct, params, IN, auxevol_gfs, stream))) return rc;\n" if has_ricci else "")
    detg = ("  if ((rc = enforce_detgammabar_constraint(rfmstruct, params, OUT, stream))) return rc;\n"
            if has_detg else "")
    sub = ("  /* substep S: IN -> OUT */\n" + ricci +
           "  if ((rc = rhs_eval_rk4_substep(S, " + rhs_call_args + ", y_n_gfs, y_nplus1_running_total_gfs, OUT, dt, "
           "stream))) return rc;\n"
           "  if ((rc = apply_bcs_curvilinear_composed(params, bcstruct, NUM_EVOL_GFS, evol_gf_parity, OUT, stream))) "
           "return rc;\n" + detg)
    steps = ""
    for s, (i, o) in enumerate((("y_n_gfs", "k_odd_gfs"), ("k_odd_gfs", "k_even_gfs"),
                                ("k_even_gfs", "k_odd_gfs"), ("k_odd_gfs", "y_n_gfs"))):
        steps += sub.replace("IN", i).replace("OUT", o).replace("substep S", "substep %d" % s) \
            .replace("(S,", "(%d," % s)
    sig = ("const rfm_struct *rfmstruct, const paramstruct *params, const bc_struct *bcstruct, double *y_n_gfs, "
           "double *k_odd_gfs, double *k_even_gfs, double *y_nplus1_running_total_gfs, double *auxevol_gfs, "
           "const double dt")
    return (
        "// One RK4 step (MoLtimestepping/C_Code_Generation.py:266-329 for RK4): four substeps, each\n"
        "// [Ricci] -> RHS + update at interior points -> ghost fill -> [det(gammabar) enforcement].\n"
        'extern "C" int MoL_step_forward_in_time(' + sig + ", hipStream_t stream) {\n"
        "  if (!rfmstruct || !params || !bcstruct || !y_n_gfs || !k_odd_gfs || !k_even_gfs ||\n"
        "      !y_nplus1_running_total_gfs) return NRPY_ERR_NULL;\n"
        + ("  if (!auxevol_gfs) return NRPY_ERR_NULL;\n" if has_ricci else "") +
        "  int rc;\n" + steps + "  return NRPY_OK;\n}\n\n"
        "struct MoL_graph {\n  hipGraph_t graph;\n  hipGraphExec_t exec;\n};\n"
        "// Capture one RK4 step into a hipGraph; replay with MoL_graph_launch. Pointers and dt are fixed\n"
        "// at capture time.\n"
        'extern "C" int MoL_graph_create(' + sig + ", MoL_graph **out) {\n"
        "  if (!out) return NRPY_ERR_NULL;\n  *out = nullptr;\n"
        "  hipStream_t s;\n"
        "  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return NRPY_ERR_LAUNCH;\n"
        "  MoL_graph *g = new MoL_graph();\n"
        "  int rc = NRPY_OK;\n"
        "  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = NRPY_ERR_LAUNCH;\n"
        "  if (!rc) {\n"
        "    rc = MoL_step_forward_in_time(rfmstruct, params, bcstruct, y_n_gfs, k_odd_gfs, k_even_gfs,\n"
        "                                  y_nplus1_running_total_gfs, auxevol_gfs, dt, s);\n"
        "    hipGraph_t gr;\n"
        "    const hipError_t e = hipStreamEndCapture(s, &gr);\n"
        "    if (!rc && e != hipSuccess) rc = NRPY_ERR_LAUNCH;\n"
        "    if (!rc) {\n"
        "      g->graph = gr;\n"
        "      if (hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0) != hipSuccess) rc = NRPY_ERR_LAUNCH;\n"
        "    }\n"
        "  }\n"
        "  (void)hipStreamDestroy(s);\n"
        "  if (rc) { delete g; return rc; }\n"
        "  *out = g;\n  return NRPY_OK;\n}\n"
        'extern "C" int MoL_graph_launch(MoL_graph *g, const int nsteps, hipStream_t stream) {\n'
        "  if (!g) return NRPY_ERR_NULL;\n  if (nsteps < 0) return NRPY_ERR_PARAM;\n"
        "  for (int n = 0; n < nsteps; n++)\n"
        "    if (hipGraphLaunch(g->exec, stream) != hipSuccess) return NRPY_ERR_LAUNCH;\n"
        "  return NRPY_OK;\n}\n"
        'extern "C" void MoL_graph_free(MoL_graph *g) {\n'
        "  if (!g) return;\n  (void)hipGraphExecDestroy(g->exec);\n  (void)hipGraphDestroy(g->graph);\n  delete g;\n}\n")


def _sommerfeld_glue(fd_order, rhs_call, shell_call=None, **defaults):
    """Sommerfeld radiation condition (CurviBoundaryConditions.py:259-658) and the RK4 step of
    the Start-to-Finish drivers' BoundaryCondition = "Sommerfeld" branch
    (Tutorial-Start_to_Finish-Solving_Maxwells_Equations_in_Vacuum-Cartesian.ipynb:240-246):
    RHS_string = rhs_eval + apply_bcs_sommerfeld on the RHS, no post-RHS boundary fill, the
    update at every point, as MoL_step_forward_in_time_sommerfeld_unfused.

    With ``shell_call`` (a call of the bundle's <fused kernel>_shell with IN, OUT, R and S
    placeholders) MoL_step_forward_in_time_sommerfeld takes the fused form where it can: per
    substep, the fused interior RHS + RK4 update (which also leaves the RHS at the interior
    points next to the boundary in the scratch R), then the radiation condition at the outer
    points fused with their update, layer by layer (apply_bcs_sommerfeld_rk4). The values are
    those of the unfused step: the same RHS and update arithmetic at every point, only the
    separate all-points update pass is gone. Grids with inner boundary points, and a step
    recorded into a graph before R exists, take the unfused step. Returns (HIP source,
    per-gridfunction parameters for the meta)."""
    from ..CurviBoundaryConditions import CurviBoundaryConditions as cbcs
    from ..MoLtimestepping import C_Code_Generation as MoL
    bcs = cbcs.sommerfeld_boundary_condition_class(fd_order=fd_order, **defaults)
    rhs_string = (rhs_call + "; apply_bcs_sommerfeld(params, xx, bcstruct, NUM_EVOL_GFS, evol_gf_parity, "
                  "RK_INPUT_GFS, RK_OUTPUT_GFS, stream);")
    sig = "const rfm_struct *rfmstruct, double *xx[3], const bc_struct *bcstruct"
    step = MoL.mol_step_function("MoL_step_forward_in_time_sommerfeld_unfused", sig, rhs_string)
    call_unfused = ("MoL_step_forward_in_time_sommerfeld_unfused(params, rfmstruct, xx, bcstruct, y_n_gfs, "
                    "k_odd_gfs, k_even_gfs, y_nplus1_running_total_gfs, dt, stream)")
    if shell_call is None:
        fused_body = "  return %s;\n" % call_unfused
    else:
        subs = ""
        for S, (i, o) in enumerate((("y_n_gfs", "k_odd_gfs"), ("k_odd_gfs", "k_even_gfs"),
                                    ("k_even_gfs", "k_odd_gfs"), ("k_odd_gfs", "y_n_gfs"))):
            subs += ("  /* substep %d: %s -> %s */\n" % (S, i, o) +
                     "  if ((rc = %s)) return rc;\n" % shell_call.replace("IN", i).replace("OUT", o)
                     .replace("SUBSTEP", str(S)) +
                     "  if ((rc = apply_bcs_sommerfeld_rk4(%d, params, xx, bcstruct, NUM_EVOL_GFS, %s, R, y_n_gfs, "
                     "y_nplus1_running_total_gfs, %s, dt, stream))) return rc;\n" % (S, i, o))
        fused_body = ("  if (!params || !bcstruct || !xx || !y_n_gfs || !k_odd_gfs || !k_even_gfs ||\n"
                      "      !y_nplus1_running_total_gfs) return NRPY_ERR_NULL;\n"
                      "  bool inner = false;\n"
                      "  for (int gz = 0; gz < NGHOSTS; gz++) inner = inner || bcstruct->num_ib_gz_pts[gz] > 0;\n"
                      "  double *R = inner ? nullptr : nrpy_sommerfeld_scratch(bcstruct, NUM_EVOL_GFS, stream);\n"
                      "  if (!R) return %s;\n  int rc;\n%s  return NRPY_OK;\n" % (call_unfused, subs))
    step += ("// One RK4 step of the Sommerfeld branch (see MoL_step_forward_in_time_sommerfeld_unfused for\n"
             "// the reference's order of operations; the fused form gives the same values).\n"
             'extern "C" int MoL_step_forward_in_time_sommerfeld(const paramstruct *params, %s, double *y_n_gfs, '
             "double *k_odd_gfs, double *k_even_gfs, double *y_nplus1_running_total_gfs, const double dt, "
             "hipStream_t stream) {\n%s}\n" % (sig, fused_body))
    meta = {"fd_order": fd_order, "at_inf": list(bcs.vars_at_infinity.values()),
            "radial_falloff_power": list(bcs.vars_radial_falloff_power.values()),
            "speed": list(bcs.vars_speed.values())}
    graph = (
        "// The Sommerfeld RK4 step captured into a hipGraph (every launch is a kernel; pointers, xx\n"
        "// and dt are fixed at capture); replay with MoL_graph_launch, free with MoL_graph_free.\n"
        'extern "C" int MoL_sommerfeld_graph_create(const paramstruct *params, const rfm_struct *rfmstruct, '
        "double *xx[3], const bc_struct *bcstruct, double *y_n_gfs, double *k_odd_gfs, double *k_even_gfs, "
        "double *y_nplus1_running_total_gfs, const double dt, MoL_graph **out) {\n"
        "  if (!out) return NRPY_ERR_NULL;\n  *out = nullptr;\n"
        "  hipStream_t s;\n"
        "  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return NRPY_ERR_LAUNCH;\n" +
        ("  if (bcstruct) {  /* the fused step's scratch, allocated before capture */\n"
         "    bool inner = false;\n"
         "    for (int gz = 0; gz < NGHOSTS; gz++) inner = inner || bcstruct->num_ib_gz_pts[gz] > 0;\n"
         "    if (!inner) (void)nrpy_sommerfeld_scratch(bcstruct, NUM_EVOL_GFS, s);\n"
         "  }\n" if shell_call is not None else "") +
        "  MoL_graph *g = new MoL_graph();\n"
        "  int rc = NRPY_OK;\n"
        "  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = NRPY_ERR_LAUNCH;\n"
        "  if (!rc) {\n"
        "    rc = MoL_step_forward_in_time_sommerfeld(params, rfmstruct, xx, bcstruct, y_n_gfs, k_odd_gfs, "
        "k_even_gfs,\n                                             y_nplus1_running_total_gfs, dt, s);\n"
        "    hipGraph_t gr;\n"
        "    const hipError_t e = hipStreamEndCapture(s, &gr);\n"
        "    if (!rc && e != hipSuccess) rc = NRPY_ERR_LAUNCH;\n"
        "    if (!rc) {\n"
        "      g->graph = gr;\n"
        "      if (hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0) != hipSuccess) rc = NRPY_ERR_LAUNCH;\n"
        "    }\n"
        "  }\n"
        "  (void)hipStreamDestroy(s);\n"
        "  if (rc) { delete g; return rc; }\n"
        "  *out = g;\n  return NRPY_OK;\n}\n")
    return ("// One RK4 step with the Sommerfeld radiation condition on the RHS.\n" + bcs.hip_source() + step + graph,
            meta)


def _mol_batch_glue(n_ricci, n_rhs):
    """MoL_batch_step / MoL_batch_graph_create: one RK4 step of every instance of a batch,
    the diagonal schedule with one launch per kernel per substep for all instances."""
    sub = ""
    for s, (i, o) in enumerate((("NRPY_BATCH_ROLE_YN", "NRPY_BATCH_ROLE_KODD"),
                                ("NRPY_BATCH_ROLE_KODD", "NRPY_BATCH_ROLE_KEVEN"),
                                ("NRPY_BATCH_ROLE_KEVEN", "NRPY_BATCH_ROLE_KODD"),
                                ("NRPY_BATCH_ROLE_KODD", "NRPY_BATCH_ROLE_YN"))):
        sub += "  /* substep %d */\n" % s
        for k in range(n_ricci):
            sub += ("  if ((rc = Ricci_eval_g%d_batched(b, %s, NRPY_BATCH_ROLE_AUXEVOL, stream))) return rc;\n"
                    % (k, i))
        for k in range(n_rhs):
            sub += ("  if ((rc = rhs_eval_rk4_substep_g%d_batched(%d, b, NRPY_BATCH_ROLE_AUXEVOL, %s, NRPY_BATCH_ROLE_YN, "
                    "NRPY_BATCH_ROLE_RT, %s, stream))) return rc;\n" % (k, s, i, o))
        sub += ("  if ((rc = nrpy_batch_apply_bcs(b, %s, NUM_EVOL_GFS, evol_gf_parity, stream))) return rc;\n"
                "  if ((rc = enforce_detgammabar_constraint_batched(b, %s, stream))) return rc;\n" % (o, o))
    return (
        "// One RK4 step of every instance of the batch (their own dt), as MoL_step_forward_in_time does\n"
        "// for one grid.\n"
        'extern "C" int MoL_batch_step(nrpy_batch *b, hipStream_t stream) {\n'
        "  if (!b) return NRPY_ERR_NULL;\n  int rc;\n" + sub + "  return NRPY_OK;\n}\n"
        "// The batch step captured into a hipGraph; replay with MoL_graph_launch, free with MoL_graph_free.\n"
        'extern "C" int MoL_batch_graph_create(nrpy_batch *b, MoL_graph **out) {\n'
        "  if (!b || !out) return NRPY_ERR_NULL;\n  *out = nullptr;\n"
        "  int rc = nrpy_batch_prepare(b);\n  if (rc) return rc;\n"
        "  hipStream_t s;\n"
        "  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return NRPY_ERR_LAUNCH;\n"
        "  MoL_graph *g = new MoL_graph();\n"
        "  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = NRPY_ERR_LAUNCH;\n"
        "  if (!rc) {\n"
        "    rc = MoL_batch_step(b, s);\n"
        "    hipGraph_t gr;\n"
        "    const hipError_t e = hipStreamEndCapture(s, &gr);\n"
        "    if (!rc && e != hipSuccess) rc = NRPY_ERR_LAUNCH;\n"
        "    if (!rc) {\n"
        "      g->graph = gr;\n"
        "      if (hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0) != hipSuccess) rc = NRPY_ERR_LAUNCH;\n"
        "    }\n"
        "  }\n"
        "  (void)hipStreamDestroy(s);\n"
        "  if (rc) { delete g; return rc; }\n"
        "  *out = g;\n  return NRPY_OK;\n}\n")


def _mol_adjoint_glue(has_rfm_vjp, state_dependent, writes_all):
    """MoL_step_adjoint: the adjoint of one RK4 step (diagonal schedule, ghost fill after
    every substep) for systems whose step is [RHS + update -> ghost fill] per substep.

    With u the interior of y_n and E the ghost fill, the step is
      k1 = R(E u), u1 = u + dt/2 k1, k2 = R(E u1), u2 = u + dt/2 k2, k3 = R(E u2),
      u3 = u + dt k3, k4 = R(E u3), u_new = u + dt/6 (k1 + 2 k2 + 2 k3 + k4),
    so for the cotangent L of u_new, with w_s = E^T R'(E u_s)^T kbar_s:
      kbar4 = dt/6 L, kbar3 = dt/3 L + dt w3, kbar2 = dt/3 L + dt/2 w2, kbar1 = dt/6 L + dt/2 w1,
      L_n = L + w3 + w2 + w1 + w0.
    When the right-hand side depends on the state (state_dependent), the stage states are
    recomputed from y_n with the forward substep kernels; for a linear right-hand side the
    adjoint does not read them. When the VJP writes every evolved gridfunction
    (writes_all), its output needs no clearing."""
    rfm_arg = "rfmstruct, " if has_rfm_vjp else ""
    vjp = (("" if writes_all else "  if ((rc = nrpy_fill(n, w, 0.0, stream))) return rc;\n") +
           "  if ((rc = rhs_eval_vjp(" + rfm_arg + "params, auxevol_gfs, STATE, kbar, vjp_G, w, nullptr, stream))) "
           "return rc;\n"
           "  if ((rc = bc_apply_T(params, bcstruct, NUM_EVOL_GFS, evol_gf_parity, w, stream))) return rc;\n")
    if state_dependent:
        recompute = (
            "  double *yn = (double *)y_n_gfs;  // substeps 0-2 read y_n only\n"
            "  if ((rc = rhs_eval_rk4_substep(0, rfmstruct, params, auxevol_gfs, yn, yn, rt, U1, dt, stream))) "
            "return rc;\n"
            "  if ((rc = apply_bcs_curvilinear_composed(params, bcstruct, NUM_EVOL_GFS, evol_gf_parity, U1, stream))) "
            "return rc;\n"
            "  if ((rc = rhs_eval_rk4_substep(1, rfmstruct, params, auxevol_gfs, U1, yn, rt, U2, dt, stream))) "
            "return rc;\n"
            "  if ((rc = apply_bcs_curvilinear_composed(params, bcstruct, NUM_EVOL_GFS, evol_gf_parity, U2, stream))) "
            "return rc;\n"
            "  if ((rc = rhs_eval_rk4_substep(2, rfmstruct, params, auxevol_gfs, U2, yn, rt, U3, dt, stream))) "
            "return rc;\n"
            "  if ((rc = apply_bcs_curvilinear_composed(params, bcstruct, NUM_EVOL_GFS, evol_gf_parity, U3, stream))) "
            "return rc;\n")
        states = ("U3", "U2", "U1", "y_n_gfs")
    else:
        recompute = "  (void)U1; (void)U2; (void)U3; (void)rt;  // linear right-hand side: the adjoint ignores the state\n"
        states = ("y_n_gfs",) * 4
    return (
        "// Adjoint of one RK4 step. On entry lambda_gfs holds the cotangent of y_{n+1} (interior points;\n"
        "// ghost cotangents zero), y_n_gfs the state at the start of the step with its ghosts filled. On\n"
        "// exit lambda_gfs holds the cotangent of y_n. U1, U2, U3, rt, acc, kbar, w are caller-allocated\n"
        "// scratch arrays of NUM_EVOL_GFS gridfunctions; vjp_G has the rhs_eval_vjp slot count.\n"
        'extern "C" int MoL_step_adjoint(const rfm_struct *rfmstruct, const paramstruct *params,\n'
        "    const bc_struct *bcstruct, const double *y_n_gfs, double *lambda_gfs, double *U1, double *U2,\n"
        "    double *U3, double *rt, double *acc, double *kbar, double *w, double *vjp_G, double *auxevol_gfs,\n"
        "    const double dt, hipStream_t stream) {\n"
        "  if (!rfmstruct || !params || !bcstruct || !y_n_gfs || !lambda_gfs || !U1 || !U2 || !U3 || !rt ||\n"
        "      !acc || !kbar || !w || !vjp_G) return NRPY_ERR_NULL;\n"
        "  int rc;\n"
        "  const size_t n = (size_t)NUM_EVOL_GFS * params->Nxx_plus_2NGHOSTS0 * params->Nxx_plus_2NGHOSTS1 *\n"
        "                   params->Nxx_plus_2NGHOSTS2;\n"
        + recompute +
        "  /* stage 4: kbar4 = dt/6 L */\n"
        "  if ((rc = nrpy_lincomb(n, kbar, dt / 6.0, lambda_gfs, 0.0, nullptr, stream))) return rc;\n"
        + vjp.replace("STATE", states[0]) +
        "  /* acc = L + w3; kbar3 = dt/3 L + dt w3 */\n"
        "  if ((rc = nrpy_adjoint_combine(n, acc, lambda_gfs, w, kbar, dt / 3.0, dt, 1, stream))) return rc;\n"
        + vjp.replace("STATE", states[1]) +
        "  if ((rc = nrpy_adjoint_combine(n, acc, lambda_gfs, w, kbar, dt / 3.0, dt / 2.0, 0, stream))) return rc;\n"
        + vjp.replace("STATE", states[2]) +
        "  if ((rc = nrpy_adjoint_combine(n, acc, lambda_gfs, w, kbar, dt / 6.0, dt / 2.0, 0, stream))) return rc;\n"
        + vjp.replace("STATE", states[3]) +
        "  /* L_n = acc + w0 */\n"
        "  return nrpy_lincomb(n, lambda_gfs, 1.0, acc, 1.0, w, stream);\n"
        "}\n")


def _fused_launcher(name, args_sig, call_args, loopopts):
    rng_interior = "InteriorPoints" in loopopts
    n = ("params->Nxx0", "params->Nxx1", "params->Nxx2") if rng_interior else \
        ("params->Nxx_plus_2NGHOSTS0", "params->Nxx_plus_2NGHOSTS1", "params->Nxx_plus_2NGHOSTS2")
    kid = bd.lp.kernel_id(name)
    block = ("dim3(%d, %d, %d)" % tuple(bd.lp.LDS_KERNELS[name]) if name in bd.lp.LDS_KERNELS else
             "nrpy_block(%d, %d, %d, %d)" % (kid, bd.lp.BLOCK_X, bd.lp.BLOCK_Y, bd.lp.BLOCK_Z))
    cases = "".join(
        "    case %d: hipLaunchKernelGGL(%s__gpu<%d>, grid, block, 0, stream, %s, y_n, rt, out, dt); break;\n"
        % (s, name, s, call_args) for s in range(4))
    return ('extern "C" int %s(const int S, %s, double *y_n, double *rt, double *out, const double dt, '
            "hipStream_t stream) {\n"
            "  if (!params || !y_n || !rt || !out) return NRPY_ERR_NULL;\n"
            "  if (S < 0 || S > 3) return NRPY_ERR_PARAM;\n"
            "  if (nrpy_check_params(params)) return NRPY_ERR_SHAPE;\n"
            "  const dim3 block = %s;\n"
            "  const dim3 grid(nrpy_nblocks(block, %s, %s, %s));\n"
            "  switch (S) {\n%s  }\n"
            "  return hipGetLastError() == hipSuccess ? 0 : NRPY_ERR_LAUNCH;\n}\n"
            % (name, args_sig, block, n[0], n[1], n[2], cases))


# the fused Sommerfeld step's interior substep in the Cartesian scalar-wave and Maxwell bundles
SHELL_CALL = ("rhs_eval_rk4_substep_shell(SUBSTEP, params, IN, y_n_gfs, y_nplus1_running_total_gfs, OUT, R, dt, "
              "stream)")


def _fused_shell_kernel(fused_kernel, name):
    """The fused RHS + RK4 kernel ``name`` (the text bd.fused_rk4_kernel wrote) as
    ``<name>_shell``: the same computation, and at the interior points next to the grid's
    boundary (one index at its first or last interior value) it also stores each RHS into
    shell_rhs, where the fused Sommerfeld step's outer-point kernels read it."""
    import re as _re
    i = fused_kernel.index("%s__gpu(" % name)
    start = fused_kernel.rindex("template <int S>", 0, i)
    end = fused_kernel.index("\n}\n", i) + 3
    k = fused_kernel[start:end].replace("%s__gpu(" % name, "%s_shell__gpu(" % name, 1)
    k = k.replace("const double dt) {", "const double dt, double *__restrict__ shell_rhs) {", 1)
    stores = _re.findall(r"NRPY_RK4_STORE\(S, \d+, __rhs_(\w+)\);", k)
    last = [m.end() for m in _re.finditer(r"NRPY_RK4_STORE\(S, \d+, __rhs_\w+\);", k)][-1]
    # points next to an i1 or i2 boundary: in the grid layout, where the i1- and i2-face kernels
    # read them. Points next to an i0 boundary: also (or only) in a compact array after the
    # NUM_EVOL_GFS grids, [gf][i2][i1][side], read by the i0-face launch; there a wave's lanes
    # and neighbouring blocks fill whole cache lines, where in the grid layout each value of
    # such a point sits on a line of its own.
    block = ("\n  if (i1 == NGHOSTS || i1 == Nxx1 + NGHOSTS - 1 || i2 == NGHOSTS || i2 == Nxx2 + NGHOSTS - 1) {\n" +
             "".join("    shell_rhs[IDX4(%sGF, i0, i1, i2)] = __rhs_%s;\n" % (g, g) for g in stores) + "  }\n"
             "  if (i0 == NGHOSTS || i0 == Nxx0 + NGHOSTS - 1) {\n"
             "    double *__restrict__ xb = shell_rhs + (size_t)NUM_EVOL_GFS * Nxx_plus_2NGHOSTS0 * "
             "Nxx_plus_2NGHOSTS1 * Nxx_plus_2NGHOSTS2 + 2 * ((size_t)i1 + (size_t)Nxx_plus_2NGHOSTS1 * i2) + "
             "(i0 == NGHOSTS ? 0 : 1);\n" +
             "".join("    xb[(size_t)%sGF * 2 * Nxx_plus_2NGHOSTS1 * Nxx_plus_2NGHOSTS2] = __rhs_%s;\n" % (g, g)
                     for g in stores) + "  }")
    return k[:last] + block + k[last:]


def _fused_shell_launcher(name, args_sig, call_args):
    """Launcher of <name>_shell, with the block shape of <name> (bd.lp.kernel_id)."""
    kid = bd.lp.kernel_id(name)
    cases = "".join(
        "    case %d: hipLaunchKernelGGL(%s_shell__gpu<%d>, grid, block, 0, stream, %s, y_n, rt, out, dt, shell_rhs); "
        "break;\n" % (s, name, s, call_args) for s in range(4))
    return ('extern "C" int %s_shell(const int S, %s, double *y_n, double *rt, double *out, double *shell_rhs, '
            "const double dt, hipStream_t stream) {\n"
            "  if (!params || !y_n || !rt || !out || !shell_rhs) return NRPY_ERR_NULL;\n"
            "  if (S < 0 || S > 3) return NRPY_ERR_PARAM;\n"
            "  if (nrpy_check_params(params)) return NRPY_ERR_SHAPE;\n"
            "  const dim3 block = nrpy_block(%d, %d, %d, %d);\n"
            "  const dim3 grid(nrpy_nblocks(block, params->Nxx0, params->Nxx1, params->Nxx2));\n"
            "  switch (S) {\n%s  }\n"
            "  return hipGetLastError() == hipSuccess ? 0 : NRPY_ERR_LAUNCH;\n}\n"
            % (name, args_sig, kid, bd.lp.BLOCK_X, bd.lp.BLOCK_Y, bd.lp.BLOCK_Z, cases))


# Output groups of the BSSN RHS, one kernel each: fewer live values per kernel. Measured on
# MI355X at 192^3 (fused RHS + RK4 update, one substep): one kernel 28 ms; two groups
# (aDD,hDD,alpha,cf,trK | lambdaU,vetU,betU) 10.2 ms; these three 8.5 ms. Grouping
# lambdaU with betU avoids evaluating the Lambdabar RHS twice (the Gamma-driver B^i RHS
# contains it).
DEFAULT_RHS_GROUPS = "aDD,trK|hDD,alpha,cf|lambdaU,vetU,betU"


def rhs_groups(names):
    """Partition RHS output names by prefix groups from NRPY_BSSN_RHS_GROUPS
    (``|``-separated groups of comma-separated name prefixes; ``all`` = one kernel)."""
    import os
    spec = os.environ.get("NRPY_BSSN_RHS_GROUPS", DEFAULT_RHS_GROUPS)
    if spec == "all":
        return [list(names)]
    groups = []
    for g in spec.split("|"):
        prefixes = [x for x in g.split(",") if x]
        sel = [n for n in names if any(n.startswith(p) for p in prefixes)]
        if sel:
            groups.append(sel)
    covered = [n for g in groups for n in g]
    if sorted(covered) != sorted(names):
        raise ValueError("RHS groups do not partition the outputs: " + spec)
    return groups


DEFAULT_RICCI_GROUPS = "RbarDD00,RbarDD01,RbarDD02|RbarDD11,RbarDD12,RbarDD22"

# LDS tile of the group kernel of the aDD and trK right-hand sides: 64x1x4, so that a
# wave covers one 64-point row for the many stencil reads that stay global. Measured at
# 192^3 (fused RK4 substep): 1.50 ms against 1.67 ms with the 16x4x4 default. The Ricci
# and Lambda-bar kernels keep 16x4x4 (with 64x1x4: 0.76 and 2.95 ms against 0.55 and
# 2.30 ms: their larger boxes take longer to copy).
ROW_TILE = (64, 1, 4)
ROW_TILE_OUTPUTS = {"aDD00", "aDD01", "aDD02", "aDD11", "aDD12", "aDD22", "trK"}
# The kernels of this output group build their x-neighbours by lane moves (codegen/lds.py):
# with them the group runs at three waves per SIMD, which hides the moves' latency
# (1.22 -> 1.03 ms per substep at 192^3). The other groups and the Ricci kernels, at one
# or two waves per SIMD, were measured slower that way (2.34 -> 2.95, 1.15 -> 1.42 ms).
XSHIFT_OUTPUTS = {"hDD00", "hDD01", "hDD02", "hDD11", "hDD12", "hDD22", "alpha", "cf"}

# The aDD,trK group needs a few registers more than two waves per SIMD allow once its
# lapse and conformal-factor stencils are differenced; asking for two waves fits it (no
# spills) and keeps its occupancy.
MIN_TWO_WAVES_OUTPUTS = {"aDD00", "aDD01", "aDD02", "aDD11", "aDD12", "aDD22", "trK"}

# Tile-relative addressing (codegen/addressing.py) of the group kernels, chosen by the
# group's outputs. Measured at 192^3, fused RK4 substep: the aDD,trK group 1.54 -> 1.44 ms
# with "xc"; the Lambda-bar group 2.34 -> 2.15 ms with "xl", but 2.66 ms with "xc" (at one
# wave per SIMD the fused 16-byte loads of x-neighbours cost it registers). Not used: the
# lane-move group with "xl" (1.03 -> 1.44 ms: its halo loads take per-lane offsets) and the
# Ricci kernels (1.16 ms either way).
TILE_ADDRESS_OUTPUTS = (
    ({"aDD00", "aDD01", "aDD02", "aDD11", "aDD12", "aDD22", "trK"}, "xc"),
    ({"betU0", "betU1", "betU2", "lambdaU0", "lambdaU1", "lambdaU2", "vetU0", "vetU1", "vetU2"}, "xl"),
)


def kernel_inputs(kernel_text):
    """Distinct (array, gridfunction) pairs a generated kernel reads: its stencil reads and
    the gridfunctions it copies into LDS."""
    pat = r"\b(in_gfs|auxevol_gfs)(?:\[IDX4\(| \+ \(size_t\)|\[\(size_t\))(\w+)GF"
    return set(re.findall(pat, kernel_text))


def lds_tile_for_group(grp):
    """The LDS tile for the kernel of an output group, or None for the default."""
    return ROW_TILE if set(grp) == ROW_TILE_OUTPUTS else None


def ricci_groups(names):
    """Partition the RbarDD components by NRPY_BSSN_RICCI_GROUPS (``|``-separated groups of
    comma-separated component names; ``all`` = one kernel)."""
    import os
    spec = os.environ.get("NRPY_BSSN_RICCI_GROUPS", DEFAULT_RICCI_GROUPS)
    if spec == "all":
        return [list(names)]
    groups = [[n for n in g.split(",") if n] for g in spec.split("|")]
    if sorted(n for g in groups for n in g) != sorted(names):
        raise ValueError("Ricci groups do not partition the components: " + spec)
    return groups


def _header(NGHOSTS):
    ps, members = bd.paramstruct_decl()
    defines, lists = bd.gf_defines_and_meta()
    parity = [bd.parity_type(n) for n in lists["evol"]]
    pre = ('#include "../nrpy_runtime_pre.h"\n#define NGHOSTS %d\n' % NGHOSTS + defines + "\n" + ps + "\n"
           + "static const int8_t evol_gf_parity[%d] = {%s};\n" % (max(1, len(parity)),
                                                                   ", ".join(str(p) for p in parity) or "0")
           + bd.check_params_code())
    meta = {"NGHOSTS": NGHOSTS, "paramstruct": members, "gfs": lists, "evol_parity": parity,
            "aux_parity": [bd.parity_type(n) for n in lists["aux"]],
            "auxevol_parity": [bd.parity_type(n) for n in lists["auxevol"]]}
    return pre, meta


def _initial_data_kernel(rhss, gfnames, loopopts="AllPoints,Read_xxs"):
    lhss = ["in_gfs[IDX4(%sGF, i0, i1, i2)]" % n.upper() for n in gfnames]
    pairs = sorted(zip(lhss, rhss), key=lambda p: p[0])
    body = oc.outputC([p[1] for p in pairs], [p[0] for p in pairs], "returnstring",
                      params="includebraces=False,outCverbose=False")
    return oc.outCfunction(outfile="returnstring", desc="Set up the initial data at all points on the numerical grid.",
                           name="initial_data",
                           params="const paramstruct *restrict params,REAL *restrict xx[3], REAL *restrict in_gfs",
                           body=body, loopopts=loopopts)


def scalarwave_cartesian():
    from ..ScalarWave import ScalarWave_RHSs as sw, InitialData as swid
    par.set_parval_from_str("grid::DIM", 3)
    par.set_parval_from_str("finite_difference::FD_CENTDERIVS_ORDER", 4)
    par.set_parval_from_str("reference_metric::CoordSystem", "Cartesian")
    rfm.reference_metric()
    sw.ScalarWave_RHSs()
    swid.InitialData(Type="PlaneWave", CoordSystem="Cartesian")
    exprs = [lhrh(lhs=gri.gfaccess("rhs_gfs", "uu"), rhs=sw.uu_rhs),
             lhrh(lhs=gri.gfaccess("rhs_gfs", "vv"), rhs=sw.vv_rhs)]
    body = fin.FD_outputC("returnstring", exprs, params="outCverbose=False")
    rhs_src = oc.outCfunction(outfile="returnstring", desc="Evaluate the scalar wave RHSs", name="rhs_eval",
                              params="const paramstruct *restrict params,const REAL *restrict in_gfs,"
                                     "REAL *restrict rhs_gfs", body=body, loopopts="InteriorPoints")
    args_decl = "const paramstruct params, const double *__restrict__ in_gfs"
    fused = bd.fused_rk4_kernel("rhs_eval_rk4_substep", args_decl, "", body, loopopts="InteriorPoints")
    fused += _fused_launcher("rhs_eval_rk4_substep", "const rfm_struct *rfmstruct, const paramstruct *params, "
                             "const double *auxevol_gfs, const double *in_gfs", "*params, in_gfs", "InteriorPoints")
    fused += _fused_shell_kernel(fused, "rhs_eval_rk4_substep") + _fused_shell_launcher(
        "rhs_eval_rk4_substep", "const paramstruct *params, const double *in_gfs", "*params, in_gfs")
    id_src = _initial_data_kernel([swid.uu_ID, swid.vv_ID], ["uu", "vv"])
    vjp_src, slots, vouts, vdep = adj.vjp_kernels("rhs_eval", exprs, "", "InteriorPoints", "")
    som_src, som_meta = _sommerfeld_glue(4, "rhs_eval(params, RK_INPUT_GFS, RK_OUTPUT_GFS, stream)",
                                         shell_call=SHELL_CALL)
    pre, meta = _header(3)
    parts = [pre, bd.coord_host_functions("Cartesian"), '#include "../nrpy_runtime.h"', bd.rfm_struct_code(),
             rhs_src, fused, id_src, vjp_src,
             _mol_glue(False, False, None, "rfmstruct, params, auxevol_gfs, IN"),
             _mol_adjoint_glue(False, vdep, sorted(vouts["EVOL"]) == sorted(gri.gridfunction_lists()[0])), som_src]
    meta["sommerfeld"] = som_meta
    meta["vjp_slots"] = [sl[0] for sl in slots]
    meta["cost"] = {"rhs_eval": bd.count_body(body)}
    meta.update({"bundle": "scalarwave_cartesian", "CoordSystem": "Cartesian", "rfm_tables": [],
                 "kernels": ["rhs_eval", "initial_data"], "has_ricci": False, "has_detg": False})
    return bd.write_bundle("scalarwave_cartesian", parts, meta)


def maxwell_cartesian():
    """Vacuum Maxwell System II in Cartesian coordinates
    (Tutorial-Start_to_Finish-Solving_Maxwells_Equations_in_Vacuum-Cartesian.ipynb):
    rhs_eval from VacuumMaxwellRHSs, toroidal-dipole initial data, the constraints
    DivE and G as a diagnostic kernel, quadratic-extrapolation outer boundaries."""
    from ..Maxwell import InitialData as mwid, VacuumMaxwell_Flat_Evol_Cartesian as mw
    par.set_parval_from_str("grid::DIM", 3)
    par.set_parval_from_str("finite_difference::FD_CENTDERIVS_ORDER", 4)
    par.set_parval_from_str("reference_metric::CoordSystem", "Cartesian")
    par.set_parval_from_str("Maxwell.InitialData::System_to_use", "System_II")
    rfm.reference_metric()
    mwid.InitialData()
    id_rhss = [mwid.AidU[0], mwid.AidU[1], mwid.AidU[2], mwid.EidU[0], mwid.EidU[1], mwid.EidU[2],
               mwid.Gamma_ID, mwid.psi_ID]
    mw.VacuumMaxwellRHSs()
    names = ["AU0", "AU1", "AU2", "EU0", "EU1", "EU2", "Gamma", "psi"]
    rhss = list(mw.ArhsU) + list(mw.ErhsU) + [mw.Gamma_rhs, mw.psi_rhs]
    exprs = [lhrh(lhs=gri.gfaccess("rhs_gfs", n), rhs=e) for n, e in zip(names, rhss)]
    body = fin.FD_outputC("returnstring", exprs, params="outCverbose=False")
    rhs_src = oc.outCfunction(outfile="returnstring", desc="Evaluate the vacuum Maxwell System II RHSs",
                              name="rhs_eval", params="const paramstruct *restrict params,"
                                                      "const REAL *restrict in_gfs,REAL *restrict rhs_gfs",
                              body=body, loopopts="InteriorPoints")
    con_body = fin.FD_outputC("returnstring", [lhrh(lhs=gri.gfaccess("aux_gfs", "DivE"), rhs=mw.C),
                                               lhrh(lhs=gri.gfaccess("aux_gfs", "G"), rhs=mw.G)],
                              params="outCverbose=False")
    con_src = oc.outCfunction(outfile="returnstring", desc="Evaluate the constraints div E and G = Gamma - div A",
                              name="Maxwell_constraints", params="const paramstruct *restrict params,"
                                                                 "const REAL *restrict in_gfs,REAL *restrict aux_gfs",
                              body=con_body, loopopts="InteriorPoints")
    args_decl = "const paramstruct params, const double *__restrict__ in_gfs"
    fused = bd.fused_rk4_kernel("rhs_eval_rk4_substep", args_decl, "", body, loopopts="InteriorPoints")
    fused += _fused_launcher("rhs_eval_rk4_substep", "const rfm_struct *rfmstruct, const paramstruct *params, "
                             "const double *auxevol_gfs, const double *in_gfs", "*params, in_gfs", "InteriorPoints")
    fused += _fused_shell_kernel(fused, "rhs_eval_rk4_substep") + _fused_shell_launcher(
        "rhs_eval_rk4_substep", "const paramstruct *params, const double *in_gfs", "*params, in_gfs")
    id_src = _initial_data_kernel(id_rhss, names)
    coords = bd.coord_host_functions("Cartesian")
    # the Start-to-Finish Maxwell run's choice: second-order radial derivative, falloff power 3
    som_src, som_meta = _sommerfeld_glue(2, "rhs_eval(params, RK_INPUT_GFS, RK_OUTPUT_GFS, stream)",
                                         shell_call=SHELL_CALL)
    pre, meta = _header(3)
    vjp_src, slots, vouts, vdep = adj.vjp_kernels("rhs_eval", exprs, "", "InteriorPoints", "")
    parts = [pre, coords, '#include "../nrpy_runtime.h"', bd.rfm_struct_code(), rhs_src, con_src, fused, id_src,
             vjp_src, _mol_glue(False, False, None, "rfmstruct, params, auxevol_gfs, IN"),
             _mol_adjoint_glue(False, vdep, sorted(vouts["EVOL"]) == sorted(gri.gridfunction_lists()[0])), som_src]
    meta["sommerfeld"] = som_meta
    meta["vjp_slots"] = [sl[0] for sl in slots]
    meta["cost"] = {"rhs_eval": bd.count_body(body)}
    meta["kernel_ids"] = dict(bd.lp.KERNEL_IDS)
    meta.update({"bundle": "maxwell_cartesian", "family": "maxwell", "CoordSystem": "Cartesian", "rfm_tables": [],
                 "kernels": ["rhs_eval", "Maxwell_constraints", "initial_data"], "has_ricci": False,
                 "has_detg": False})
    return bd.write_bundle("maxwell_cartesian", parts, meta)


def scalarwave_curvilinear(CoordSystem, bundle_name, NGHOSTS=3):
    """Scalar wave in a curvilinear reference metric with precomputed rfm tables
    (Tutorial-Start_to_Finish-ScalarWaveCurvilinear.ipynb: ScalarWaveCurvilinear_RHSs, plane-wave
    data, quadratic extrapolation + parity boundary conditions). NGHOSTS=3 is the ghost depth of
    the reference's curvilinear boundary-condition validation (FD4 + 1)."""
    from ..ScalarWave import ScalarWaveCurvilinear_RHSs as sw, InitialData as swid
    par.set_parval_from_str("grid::DIM", 3)
    par.set_parval_from_str("finite_difference::FD_CENTDERIVS_ORDER", 4)
    par.set_parval_from_str("reference_metric::CoordSystem", CoordSystem)
    rfm.reference_metric()
    swid.InitialData(Type="PlaneWave", CoordSystem=CoordSystem)
    id_rhss = [swid.uu_ID, swid.vv_ID]
    par.set_parval_from_str("reference_metric::enable_rfm_precompute", "True")
    rfm.reference_metric()
    sw.ScalarWaveCurvilinear_RHSs()
    exprs = [lhrh(lhs=gri.gfaccess("rhs_gfs", "uu"), rhs=sw.uu_rhs),
             lhrh(lhs=gri.gfaccess("rhs_gfs", "vv"), rhs=sw.vv_rhs)]
    body = fin.FD_outputC("returnstring", exprs, params="outCverbose=False")
    rhs_src = oc.outCfunction(outfile="returnstring", desc="Evaluate the scalar wave RHSs, curvilinear coordinates",
                              name="rhs_eval",
                              params="rfm_struct *restrict rfmstruct,const paramstruct *restrict params,"
                                     "const REAL *restrict in_gfs,REAL *restrict rhs_gfs",
                              body=body, loopopts="InteriorPoints,Enable_rfm_precompute")
    args_decl = ("const paramstruct params, const rfm_struct rfmstruct, const double *__restrict__ auxevol_gfs, "
                 "const double *__restrict__ in_gfs")
    fused = bd.fused_rk4_kernel("rhs_eval_rk4_substep", args_decl, "", body,
                                loopopts="InteriorPoints,Enable_rfm_precompute")
    fused += _fused_launcher("rhs_eval_rk4_substep", "const rfm_struct *rfmstruct, const paramstruct *params, "
                             "const double *auxevol_gfs, const double *in_gfs",
                             "*params, *rfmstruct, auxevol_gfs, in_gfs", "InteriorPoints")
    rfm_code = bd.rfm_struct_code()
    tables = [(n, str(e), list(a)) for n, e, a in rfm.rfm_precompute_tables()]
    par.set_parval_from_str("reference_metric::enable_rfm_precompute", "False")
    rfm.ref_metric__hatted_quantities()
    id_src = _initial_data_kernel(id_rhss, ["uu", "vv"])
    # The EigenCoord's reference metric registers its own C parameters (e.g. RMAX), so
    # the coordinate functions are generated before the paramstruct is declared.
    coords = bd.coord_host_functions(CoordSystem)
    som_src, som_meta = _sommerfeld_glue(4, "rhs_eval(rfmstruct, params, RK_INPUT_GFS, RK_OUTPUT_GFS, stream)")
    pre, meta = _header(NGHOSTS)
    parts = [pre, coords, '#include "../nrpy_runtime.h"', rfm_code,
             rhs_src, fused, id_src, _mol_glue(False, False, None, "rfmstruct, params, auxevol_gfs, IN"), som_src]
    meta["sommerfeld"] = som_meta
    meta["cost"] = {"rhs_eval": bd.count_body(body)}
    meta["kernel_ids"] = dict(bd.lp.KERNEL_IDS)
    meta.update({"bundle": bundle_name, "family": "scalarwave_curvilinear", "CoordSystem": CoordSystem,
                 "rfm_tables": tables,
                 "kernels": ["rhs_eval", "initial_data"], "has_ricci": False, "has_detg": False})
    return bd.write_bundle(bundle_name, parts, meta)


def maxwell_curvilinear(CoordSystem, bundle_name, NGHOSTS=3):
    """Vacuum Maxwell System II in a curvilinear reference metric with rescaled vectors, the
    setup of Tutorial-Start_to_Finish-Solving_Maxwells_Equations_in_Vacuum-Curvilinear.ipynb:
    rhs_eval from VacuumMaxwellRHSs_rescaled with rfm precompute tables, toroidal-dipole data
    A^i/ReU^i, E^i/ReU^i at time = params->time (the exact solution at that time), the
    constraints C = Dhat_i E^i and G, the Sommerfeld condition with fd_order 4, falloff power
    3, speed 1 and value 0 at infinity (the driver's BoundaryCondition = "Sommerfeld"), and
    the quadratic-extrapolation step as the alternative."""
    from ..Maxwell import InitialData as mwid, VacuumMaxwell_Flat_Evol_Curvilinear_rescaled as mw
    par.set_parval_from_str("grid::DIM", 3)
    par.set_parval_from_str("finite_difference::FD_CENTDERIVS_ORDER", 4)
    par.set_parval_from_str("reference_metric::CoordSystem", CoordSystem)
    par.set_parval_from_str("Maxwell.InitialData::System_to_use", "System_II")
    rfm.reference_metric()
    mwid.InitialData()
    id_names = ["aU0", "aU1", "aU2", "eU0", "eU1", "eU2", "psi", "Gamma"]
    id_rhss = ([mwid.AidU[i] / rfm.ReU[i] for i in range(3)] + [mwid.EidU[i] / rfm.ReU[i] for i in range(3)]
               + [mwid.psi_ID, mwid.Gamma_ID])
    par.set_parval_from_str("reference_metric::enable_rfm_precompute", "True")
    rfm.reference_metric()
    mw.VacuumMaxwellRHSs_rescaled()
    names = ["aU0", "aU1", "aU2", "eU0", "eU1", "eU2", "psi", "Gamma"]
    rhss = list(mw.arhsU) + list(mw.erhsU) + [mw.psi_rhs, mw.Gamma_rhs]
    exprs = [lhrh(lhs=gri.gfaccess("rhs_gfs", n), rhs=e) for n, e in zip(names, rhss)]
    body = fin.FD_outputC("returnstring", exprs, params="outCverbose=False")
    rhs_src = oc.outCfunction(outfile="returnstring",
                              desc="Evaluate the vacuum Maxwell System II RHSs, rescaled curvilinear form",
                              name="rhs_eval",
                              params="rfm_struct *restrict rfmstruct,const paramstruct *restrict params,"
                                     "const REAL *restrict in_gfs,REAL *restrict rhs_gfs",
                              body=body, loopopts="InteriorPoints,Enable_rfm_precompute")
    gri.register_gridfunctions("AUX", ["C", "G"])
    con_body = fin.FD_outputC("returnstring", [lhrh(lhs=gri.gfaccess("aux_gfs", "C"), rhs=mw.C),
                                               lhrh(lhs=gri.gfaccess("aux_gfs", "G"), rhs=mw.G)],
                              params="outCverbose=False")
    con_src = oc.outCfunction(outfile="returnstring", desc="Evaluate the constraints C = Dhat_i E^i and G",
                              name="Maxwell_constraints",
                              params="rfm_struct *restrict rfmstruct,const paramstruct *restrict params,"
                                     "const REAL *restrict in_gfs,REAL *restrict aux_gfs",
                              body=con_body, loopopts="InteriorPoints,Enable_rfm_precompute")
    args_decl = ("const paramstruct params, const rfm_struct rfmstruct, const double *__restrict__ auxevol_gfs, "
                 "const double *__restrict__ in_gfs")
    fused = bd.fused_rk4_kernel("rhs_eval_rk4_substep", args_decl, "", body,
                                loopopts="InteriorPoints,Enable_rfm_precompute")
    fused += _fused_launcher("rhs_eval_rk4_substep", "const rfm_struct *rfmstruct, const paramstruct *params, "
                             "const double *auxevol_gfs, const double *in_gfs",
                             "*params, *rfmstruct, auxevol_gfs, in_gfs", "InteriorPoints")
    rfm_code = bd.rfm_struct_code()
    tables = [(n, str(e), list(a)) for n, e, a in rfm.rfm_precompute_tables()]
    par.set_parval_from_str("reference_metric::enable_rfm_precompute", "False")
    rfm.ref_metric__hatted_quantities()
    id_src = _initial_data_kernel(id_rhss, id_names)
    coords = bd.coord_host_functions(CoordSystem)
    som_src, som_meta = _sommerfeld_glue(4, "rhs_eval(rfmstruct, params, RK_INPUT_GFS, RK_OUTPUT_GFS, stream)")
    pre, meta = _header(NGHOSTS)
    parts = [pre, coords, '#include "../nrpy_runtime.h"', rfm_code, rhs_src, con_src, fused, id_src,
             _mol_glue(False, False, None, "rfmstruct, params, auxevol_gfs, IN"), som_src]
    meta["sommerfeld"] = som_meta
    meta["cost"] = {"rhs_eval": bd.count_body(body)}
    meta["kernel_ids"] = dict(bd.lp.KERNEL_IDS)
    meta.update({"bundle": bundle_name, "family": "maxwell_curvilinear", "CoordSystem": CoordSystem,
                 "rfm_tables": tables, "kernels": ["rhs_eval", "Maxwell_constraints", "initial_data"],
                 "has_ricci": False, "has_detg": False})
    return bd.write_bundle(bundle_name, parts, meta)


BSSN_DIFFERENCED_GFS = "alpha,cf"


def VJP_GROUPS(groups):
    """Output groups of the BSSN RHS VJP's pass 1: the forward output groups, or the
    prefix groups of NRPY_BSSN_VJP_GROUPS (measurements; same syntax as
    NRPY_BSSN_RHS_GROUPS). Finer groups spill less but recompute the shared forward
    temporaries and add to each other's slots: at 96^3 one adjoint RK4 step took 38.6 ms
    with the forward groups, 42.0 ms with the Lambda-bar group split into betU,vetU and
    lambdaU, 41.4 ms into betU, vetU and lambdaU, 48.5 ms with lambdaU per component."""
    import os
    spec = os.environ.get("NRPY_BSSN_VJP_GROUPS")
    if not spec:
        return [list(grp) for grp in groups]
    names = [n for grp in groups for n in grp]
    out = [[n for n in names if any(n.startswith(p) for p in g.split(",") if p)] for g in spec.split("|")]
    out = [g for g in out if g]
    if sorted(n for g in out for n in g) != sorted(names):
        raise ValueError("VJP groups do not partition the outputs: " + spec)
    return out


def bssn(CoordSystem, symmetry_axes, bundle_name, FD_order=4):
    from ..BSSN import BSSN_quantities as Bq, BSSN_RHSs as rhs, BSSN_gauge_RHSs as gaugerhs
    from ..BSSN import Enforce_Detgammabar_Constraint as EGC, BSSN_constraints as bssncon
    from ..BSSN import BrillLindquist as bl
    t0 = time.time()
    par.set_parval_from_str("grid::DIM", 3)
    par.set_parval_from_str("finite_difference::FD_CENTDERIVS_ORDER", FD_order)
    par.set_parval_from_str("reference_metric::CoordSystem", CoordSystem)
    par.set_parval_from_str("indexedexp::symmetry_axes", symmetry_axes)
    # The lapse and the conformal factor are ~1: their stencils work on differences from the
    # centre value, so their second derivatives keep full precision (FD_difference_from_centre).
    par.set_parval_from_str("finite_difference::FD_difference_from_centre", BSSN_DIFFERENCED_GFS)
    rfm.reference_metric()
    # Initial data is written in closed form (Read_xxs), before rfm precompute is switched on.
    bl.BrillLindquist()
    id_rhss = [bl.trK, bl.alpha, bl.cf] + [v for i in range(3) for v in (bl.lambdaU[i], bl.vetU[i], bl.betU[i])] + \
        [v for i in range(3) for j in range(i, 3) for v in (bl.hDD[i][j], bl.aDD[i][j])]
    id_names = ["trK", "alpha", "cf"] + [v for i in range(3) for v in ("lambdaU%d" % i, "vetU%d" % i, "betU%d" % i)] + \
        [v for i in range(3) for j in range(i, 3) for v in ("hDD%d%d" % (i, j), "aDD%d%d" % (i, j))]

    par.set_parval_from_str("BSSN.BSSN_gauge_RHSs::ShiftEvolutionOption", "GammaDriving2ndOrder_Covariant")
    par.set_parval_from_str("reference_metric::enable_rfm_precompute", "True")
    rfm.reference_metric()
    par.set_parval_from_str("BSSN.BSSN_quantities::LeaveRicciSymbolic", "True")
    rhs.BSSN_RHSs()
    gaugerhs.BSSN_gauge_RHSs()
    Bq.BSSN_basic_tensors()
    betaU = Bq.betaU
    detg_exprs = EGC.Enforce_Detgammabar_Constraint_symb_expressions()
    par.set_parval_from_str("BSSN.BSSN_quantities::LeaveRicciSymbolic", "False")
    Bq.RicciBar__gammabarDD_dHatD__DGammaUDD__DGammaU()
    RbarDD = Bq.RbarDD
    gri.register_gridfunctions("AUX", "H")
    ixp.register_gridfunctions_for_single_rank1("AUX", "MU")
    bssncon.BSSN_constraints(add_T4UUmunu_source_terms=False)
    print("(BENCH) symbolic expressions %.1fs" % (time.time() - t0), flush=True)

    lhs_names = ["alpha", "cf", "trK"]
    rhs_exprs = [gaugerhs.alpha_rhs, rhs.cf_rhs, rhs.trK_rhs]
    for i in range(3):
        lhs_names += ["betU%d" % i, "lambdaU%d" % i, "vetU%d" % i]
        rhs_exprs += [gaugerhs.bet_rhsU[i], rhs.lambda_rhsU[i], gaugerhs.vet_rhsU[i]]
        for j in range(i, 3):
            lhs_names += ["aDD%d%d" % (i, j), "hDD%d%d" % (i, j)]
            rhs_exprs += [rhs.a_rhsDD[i][j], rhs.h_rhsDD[i][j]]
    pairs = sorted(zip(lhs_names, rhs_exprs), key=lambda p: p[0])
    evol_rhss = [lhrh(lhs=gri.gfaccess("rhs_gfs", n), rhs=e) for n, e in pairs]
    t0 = time.time()
    groups = rhs_groups([n for n, _ in pairs])
    rhs_params = ("rfm_struct *restrict rfmstruct,const paramstruct *restrict params,"
                  "const REAL *restrict auxevol_gfs,const REAL *restrict in_gfs,REAL *restrict rhs_gfs")
    args_decl = ("const paramstruct params, const rfm_struct rfmstruct, const double *__restrict__ auxevol_gfs, "
                 "const double *__restrict__ in_gfs")
    rhs_src, fused, group_costs, group_inputs = "", "", [], []
    # Large Cartesian grids: the RHS, Ricci and Hamiltonian kernels read their
    # mixed-derivative stencils from LDS tiles (codegen/lds.py). The small
    # curvilinear grids of the Brill-Lindquist runs (two points in phi) keep the
    # plain kernels.
    bd.lp.LDS_STAGING = (FD_order // 2 + 1) if CoordSystem == "Cartesian" else 0
    lds.MIN_NXX0 = 0
    lds.XSHIFT_KERNELS = set()
    addressing.TILE_ADDRESSING = {}
    addressing.TILE_BOUND[:] = [0, 0, 0]
    addressing.APPLIED.clear()
    addressing.NGHOSTS = FD_order // 2 + 1
    for k, grp in enumerate(groups):
        if CoordSystem == "Cartesian" and set(grp) == MIN_TWO_WAVES_OUTPUTS:
            bd.lp.MIN_WAVES_PER_EU["rhs_eval_rk4_substep_g%d" % k] = bd.lp.MIN_WAVES_PER_EU["rhs_eval_g%d" % k] = 2
        if bd.lp.LDS_STAGING and set(grp) == XSHIFT_OUTPUTS:
            lds.XSHIFT_KERNELS |= {"rhs_eval_g%d" % k, "rhs_eval_rk4_substep_g%d" % k}
        for outs, form in TILE_ADDRESS_OUTPUTS if bd.lp.LDS_STAGING else ():
            if set(grp) == outs:
                addressing.TILE_ADDRESSING["rhs_eval_g%d" % k] = form
                addressing.TILE_ADDRESSING["rhs_eval_rk4_substep_g%d" % k] = form
        if bd.lp.LDS_STAGING and lds_tile_for_group(grp):
            bd.lp.LDS_TILES["rhs_eval_g%d" % k] = bd.lp.LDS_TILES["rhs_eval_rk4_substep_g%d" % k] = \
                lds_tile_for_group(grp)
        sel = [e for (n, _), e in zip(pairs, evol_rhss) if n in grp]
        body = fin.FD_outputC("returnstring", sel, params="outCverbose=False", upwindcontrolvec=betaU)
        group_costs.append(bd.count_body(body))
        rhs_src += oc.outCfunction(outfile="returnstring", desc="BSSN RHSs of %s" % ", ".join(grp),
                                   name="rhs_eval_g%d" % k, params=rhs_params, body=body,
                                   loopopts="InteriorPoints,Enable_rfm_precompute")
        kern = bd.fused_rk4_kernel("rhs_eval_rk4_substep_g%d" % k, args_decl, "", body,
                                   loopopts="InteriorPoints,Enable_rfm_precompute")
        group_inputs.append(len(kernel_inputs(kern)))
        fused += kern
        fused += _fused_launcher("rhs_eval_rk4_substep_g%d" % k, "const rfm_struct *rfmstruct, "
                                 "const paramstruct *params, const double *auxevol_gfs, const double *in_gfs",
                                 "*params, *rfmstruct, auxevol_gfs, in_gfs", "InteriorPoints")
    print("(BENCH) rhs_eval codegen %.1fs (%d groups)" % (time.time() - t0, len(groups)), flush=True)
    calls = "".join("  if ((rc = rhs_eval_g%d(rfmstruct, params, auxevol_gfs, in_gfs, rhs_gfs, stream))) return rc;\n"
                    % k for k in range(len(groups)))
    rhs_src += ("/* Evaluate the BSSN RHSs: the output gridfunctions are computed by %d kernels, each with\n"
                " * the derivatives and intermediates its outputs need, to bound register pressure. */\n"
                'extern "C" int rhs_eval(const rfm_struct *rfmstruct, const paramstruct *params, '
                "const double *auxevol_gfs, const double *in_gfs, double *rhs_gfs, hipStream_t stream) {\n"
                "  int rc;\n%s  return NRPY_OK;\n}\n" % (len(groups), calls))
    calls = "".join("  if ((rc = rhs_eval_rk4_substep_g%d(S, rfmstruct, params, auxevol_gfs, in_gfs, y_n, rt, out, "
                    "dt, stream))) return rc;\n" % k for k in range(len(groups)))
    fused += ('extern "C" int rhs_eval_rk4_substep(const int S, const rfm_struct *rfmstruct, const paramstruct *params, '
              "const double *auxevol_gfs, const double *in_gfs, double *y_n, double *rt, double *out, "
              "const double dt, hipStream_t stream) {\n  int rc;\n%s  return NRPY_OK;\n}\n" % calls)
    t0 = time.time()
    ric_names = ["RbarDD%d%d" % (i, j) for i in range(3) for j in range(i, 3)]
    ric_exprs = {"RbarDD%d%d" % (i, j): RbarDD[i][j] for i in range(3) for j in range(i, 3)}
    ric_groups = ricci_groups(ric_names)
    ricci_params = ("rfm_struct *restrict rfmstruct,const paramstruct *restrict params,"
                    "const REAL *restrict in_gfs,REAL *restrict auxevol_gfs")
    ricci_src, ricci_costs = "", []
    for k, grp in enumerate(ric_groups):
        body_k = fin.FD_outputC("returnstring", [lhrh(lhs=gri.gfaccess("auxevol_gfs", n), rhs=ric_exprs[n])
                                                 for n in grp], params="outCverbose=False")
        ricci_costs.append(bd.count_body(body_k))
        ricci_src += oc.outCfunction(outfile="returnstring", desc="Evaluate the Ricci tensor components %s"
                                     % ", ".join(grp), name="Ricci_eval_g%d" % k, params=ricci_params, body=body_k,
                                     loopopts="InteriorPoints,Enable_rfm_precompute")
    calls = "".join("  if ((rc = Ricci_eval_g%d(rfmstruct, params, in_gfs, auxevol_gfs, stream))) return rc;\n" % k
                    for k in range(len(ric_groups)))
    ricci_src += ("/* Evaluate the Ricci tensor (%d kernels, one per group of components). */\n"
                  'extern "C" int Ricci_eval(const rfm_struct *rfmstruct, const paramstruct *params, '
                  "const double *in_gfs, double *auxevol_gfs, hipStream_t stream) {\n"
                  "  int rc;\n%s  return NRPY_OK;\n}\n" % (len(ric_groups), calls))
    print("(BENCH) Ricci codegen %.1fs (%d groups)" % (time.time() - t0, len(ric_groups)), flush=True)
    t0 = time.time()
    h_body = fin.FD_outputC("returnstring", lhrh(lhs=gri.gfaccess("aux_gfs", "H"), rhs=bssncon.H),
                            params="outCverbose=False")
    print("(BENCH) Hamiltonian codegen %.1fs" % (time.time() - t0), flush=True)
    h_src = oc.outCfunction(outfile="returnstring", desc="Evaluate the Hamiltonian constraint",
                            name="Hamiltonian_constraint",
                            params="rfm_struct *restrict rfmstruct,const paramstruct *restrict params,"
                                   "const REAL *restrict in_gfs, REAL *restrict aux_gfs",
                            body=h_body, loopopts="InteriorPoints,Enable_rfm_precompute")
    # momentum constraint, as the reference's exact-initial-data driver emits it
    # (Tutorial-Start_to_Finish-BSSNCurvilinear-Setting_up_Exact_Initial_Data.ipynb, cell 17)
    mu_body = fin.FD_outputC("returnstring", [lhrh(lhs=gri.gfaccess("aux_gfs", "MU%d" % i), rhs=bssncon.MU[i])
                                              for i in range(3)], params="outCverbose=False")
    mu_src = oc.outCfunction(outfile="returnstring", desc="Evaluate the momentum constraint",
                             name="momentum_constraint",
                             params="rfm_struct *restrict rfmstruct,const paramstruct *restrict params,"
                                    "const REAL *restrict in_gfs, REAL *restrict aux_gfs",
                             body=mu_body, loopopts="InteriorPoints,Enable_rfm_precompute")
    bd.lp.LDS_STAGING = 0
    detg_body = fin.FD_outputC("returnstring", detg_exprs, params="outCverbose=False")
    detg_src = oc.outCfunction(outfile="returnstring", desc="Enforce det(gammabar) = det(gammahat) constraint.",
                               name="enforce_detgammabar_constraint",
                               params="const rfm_struct *restrict rfmstruct,const paramstruct *restrict params, "
                                      "REAL *restrict in_gfs",
                               body=detg_body, loopopts="AllPoints,Enable_rfm_precompute")
    # Precompute tables depend on which functions remain; build the struct now.
    rfm_code = bd.rfm_struct_code()
    tables = [(n, str(e), list(a)) for n, e, a in rfm.rfm_precompute_tables()]
    par.set_parval_from_str("reference_metric::enable_rfm_precompute", "False")
    rfm.ref_metric__hatted_quantities()
    t0 = time.time()
    id_src = _initial_data_kernel(id_rhss, id_names)
    # d(initial data)/d(BH1_mass), d/d(BH2_mass) at all points
    dm_exprs, dm_names = [], []
    for arr, m in (("dID_dm1", bl.BH1_mass), ("dID_dm2", bl.BH2_mass)):
        for n, e in sorted(zip(id_names, id_rhss)):
            dm_exprs.append(sp.diff(e, m))
            dm_names.append("%s[IDX4(%sGF, i0, i1, i2)]" % (arr, n.upper()))
    dm_body = oc.outputC(dm_exprs, dm_names, "returnstring", params="includebraces=False,outCverbose=False")
    id_dmass_src = oc.outCfunction(outfile="returnstring",
                                   desc="Derivatives of the Brill-Lindquist initial data with respect to the masses.",
                                   name="initial_data_dmass",
                                   params="const paramstruct *restrict params,REAL *restrict xx[3], "
                                          "REAL *restrict dID_dm1, REAL *restrict dID_dm2",
                                   body=dm_body, loopopts="AllPoints,Read_xxs")
    print("(BENCH) initial data codegen %.1fs" % (time.time() - t0), flush=True)
    # Vector-Jacobian products of the whole RHS (all outputs, inputs in_gfs and RbarDD) and of
    # the Ricci tensor (inputs in_gfs), for adjoints through rhs_eval and Ricci_eval.
    t0 = time.time()
    # Pass 1 runs per output group, as the forward kernels do; pass 2 per cotangent gridfunction.
    # Tile-relative addressing (Cartesian grids), measured on one 96^3 adjoint RK4 step: 38.5 ms
    # without; 37.8 with the transposes in form "xc" (38.3 in "xl"); 36.5 with the RHS pass-1
    # kernels in "xl" as well (41.7 in "xc"); 37.2 with the Ricci pass-1 kernels in "xl" too.
    if CoordSystem == "Cartesian":
        addressing.TILE_ADDRESSING.update({"rhs_eval_vjp_transpose_*": "xc", "Ricci_eval_vjp_transpose_*": "xc",
                                           "rhs_eval_vjp_pointwise_*": "xl"})
    vjp_groups = VJP_GROUPS(groups)
    vjp_src, vslots, _, _ = adj.vjp_kernels("rhs_eval", evol_rhss, betaU, "InteriorPoints,Enable_rfm_precompute",
                                            "rfm_struct *restrict rfmstruct,",
                                            groups=[[i for i, (n, _) in enumerate(pairs) if n in grp]
                                                    for grp in vjp_groups], split_transpose=True)
    rvjp_src, rslots, _, _ = adj.vjp_kernels(
        "Ricci_eval", [lhrh(lhs=gri.gfaccess("auxevol_gfs", n), rhs=ric_exprs[n]) for n in ric_names], "",
        "InteriorPoints,Enable_rfm_precompute", "rfm_struct *restrict rfmstruct,",
        groups=[[ric_names.index(n) for n in grp] for grp in ric_groups], split_transpose=True)
    # det(gammabar) enforcement and the Hamiltonian constraint, for the adjoint of the BSSN RK4
    # step and of the H-L2 diagnostic; the mass derivatives of the initial data.
    dvjp_src, dslots, _, _ = adj.vjp_kernels("enforce_detgammabar_constraint", detg_exprs, "",
                                             "AllPoints,Enable_rfm_precompute", "rfm_struct *restrict rfmstruct,",
                                             all_points=True)
    hvjp_src, hslots, _, _ = adj.vjp_kernels("Hamiltonian_constraint", [lhrh(lhs=gri.gfaccess("aux_gfs", "H"),
                                                                             rhs=bssncon.H)], "",
                                             "InteriorPoints,Enable_rfm_precompute", "rfm_struct *restrict rfmstruct,")
    print("(BENCH) VJP codegen %.1fs (%d + %d + %d + %d slots)" % (time.time() - t0, len(vslots), len(rslots),
                                                                  len(dslots), len(hslots)), flush=True)
    batch_names = (["Ricci_eval_g%d" % k for k in range(len(ric_groups))] +
                   ["rhs_eval_rk4_substep_g%d" % k for k in range(len(groups))] + ["enforce_detgammabar_constraint"])
    batch_src, batch_table = bt.batched_kernels(batch_names)
    pre, meta = _header(FD_order // 2 + 1)
    parts = [pre, bd.coord_host_functions(CoordSystem), '#include "../nrpy_runtime.h"', rfm_code,
             ricci_src, rhs_src, fused, h_src, mu_src, detg_src, id_src, id_dmass_src, vjp_src, rvjp_src, dvjp_src,
             hvjp_src,
             _mol_glue(True, True, None, "rfmstruct, params, auxevol_gfs, IN"),
             bt.instance_struct(), batch_table, '#include "../nrpy_batch.h"', batch_src,
             _mol_batch_glue(len(ric_groups), len(groups))]
    meta["rhs_groups"] = groups
    # distinct gridfunctions each group's kernel reads (evolved and RbarDD): its algorithmic
    # traffic is 8 (inputs + 4 outputs) bytes per point
    meta["rhs_group_inputs"] = group_inputs
    meta["ricci_groups"] = ric_groups
    meta["vjp_slots"] = [sl[0] for sl in vslots]
    meta["ricci_vjp_slots"] = [sl[0] for sl in rslots]
    meta["detg_vjp_slots"] = [sl[0] for sl in dslots]
    meta["hamiltonian_vjp_slots"] = [sl[0] for sl in hslots]
    meta["kernel_ids"] = dict(bd.lp.KERNEL_IDS)
    meta["lds_kernels"] = dict(bd.lp.LDS_KERNELS)
    meta["xshift_kernels"] = sorted(lds.XSHIFT_KERNELS)
    meta["tile_addressing"] = dict(sorted(addressing.APPLIED.items()))
    meta["cost"] = {"rhs_eval": {"flops": sum(c["flops"] for c in group_costs), "groups": group_costs},
                    "Ricci_eval": {"flops": sum(c["flops"] for c in ricci_costs), "groups": ricci_costs},
                    "Hamiltonian_constraint": bd.count_body(h_body),
                    "momentum_constraint": bd.count_body(mu_body),
                    "enforce_detgammabar_constraint": bd.count_body(detg_body)}
    meta.update({"bundle": bundle_name, "CoordSystem": CoordSystem, "symmetry_axes": symmetry_axes,
                 "rfm_tables": tables, "has_ricci": True, "has_detg": True,
                 "kernels": ["Ricci_eval", "rhs_eval", "Hamiltonian_constraint", "momentum_constraint",
                             "enforce_detgammabar_constraint",
                             "initial_data"]})
    return bd.write_bundle(bundle_name, parts, meta)


# Launch shape of the memory-bound Cartesian bundles. Block-shape sweep of the fused
# RHS + RK4 substep on MI355X: scalar wave 256^3 0.310 ms with 64x4x1 against 0.329 ms
# with the BSSN default 64x1x4; vacuum Maxwell 192^3 0.585 against 0.616 ms.
MEMORY_BOUND_BLOCK = {"scalarwave_cartesian": (64, 4, 1), "maxwell_cartesian": (64, 4, 1)}


def generate(name):
    saved = (bd.lp.BLOCK_X, bd.lp.BLOCK_Y, bd.lp.BLOCK_Z)
    if name in MEMORY_BOUND_BLOCK:
        bd.lp.BLOCK_X, bd.lp.BLOCK_Y, bd.lp.BLOCK_Z = MEMORY_BOUND_BLOCK[name]
    try:
        return _generate(name)
    finally:
        bd.lp.BLOCK_X, bd.lp.BLOCK_Y, bd.lp.BLOCK_Z = saved


def _generate(name):
    if name == "scalarwave_cartesian":
        return scalarwave_cartesian()
    if name == "bssn_cartesian":
        return bssn("Cartesian", "", name)
    if name == "bssn_spherical_sym2":
        return bssn("Spherical", "2", name)
    if name in CURVI_SCALARWAVE:
        return scalarwave_curvilinear(CURVI_SCALARWAVE[name], name)
    if name == "maxwell_cartesian":
        return maxwell_cartesian()
    if name in CURVI_MAXWELL:
        return maxwell_curvilinear(CURVI_MAXWELL[name], name)
    if name == "bssn_etk":
        from .etk import bssn_etk
        return bssn_etk(name)
    raise ValueError("unknown bundle " + name)


if __name__ == "__main__":
    for n in sys.argv[1:]:
        t = time.time()
        print("generated", generate(n), "in %.1fs" % (time.time() - t), flush=True)
