"""Symbolic expression trees vs the reference's trusted values.

Each case mirrors one entry of the reference's unit tests: the same function
call after the same initialisation (BSSN/tests/test_BSSN.py,
ScalarWave/tests/test_ScalarWave.py, Maxwell/tests/test_Maxwell.py,
tests/test_reference_metric.py), evaluated with the md5-seeded scheme of
UnitTesting/ and compared with the values recorded in the reference's
trusted_values_dict.py files (extracted to tests/golden/trusted_values.json by
tests/golden/extract_trusted_values.py). As in the reference, every case runs
in a fresh interpreter so that parameter and gridfunction registries start
empty.
"""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

RFM_NAMES = ['xxmin', 'xxmax', 'UnitVectors', 'ReU', 'ReDD', 'ghatDD', 'ghatUU', 'detgammahat',
             'detgammahatdD', 'detgammahatdDD', 'ReUdD', 'ReUdDD', 'ReDDdD', 'ReDDdDD', 'ghatDDdD',
             'ghatDDdDD', 'GammahatUDD', 'GammahatUDDdD', 'Cart_to_xx', 'xx_to_Cart', 'xxSph', 'scalefactor_orthog']

RFM_INIT = "rfm.reference_metric()\n"

# (trusted entry, setup code, module expression, globals)
CASES = []
for cs in ["Cartesian", "Spherical", "SinhSpherical", "SinhSphericalv2", "Cylindrical", "SinhCylindrical",
           "SinhCylindricalv2", "SymTP", "SinhSymTP"]:
    CASES.append(("rfm_%s__reference_metric__True__globals" % cs,
                  'par.set_parval_from_str("reference_metric::CoordSystem", "%s")\nrfm.reference_metric(True)\n' % cs,
                  "rfm", RFM_NAMES))

CASES += [
    ("ScalarWave_RHSs__ScalarWave_RHSs__globals",
     "from nrpytutorial_amd.ScalarWave import ScalarWave_RHSs as m\nm.ScalarWave_RHSs()\n"
     "from nrpytutorial_amd.ScalarWave import CommonParams\nm.wavespeed = CommonParams.wavespeed\n",
     "m", ["wavespeed", "uu_rhs", "vv_rhs"]),
    ("ScalarWaveCurvilinear_RHSs__ScalarWaveCurvilinear_RHSs__globals",
     "from nrpytutorial_amd.ScalarWave import ScalarWaveCurvilinear_RHSs as m\nm.ScalarWaveCurvilinear_RHSs()\n",
     "m", ["uu_rhs", "vv_rhs"]),
    ("InitialData__InitialData__Type__PlaneWave___globals",
     "from nrpytutorial_amd.ScalarWave import InitialData as m\nm.InitialData(Type='PlaneWave')\n",
     "m", ["uu_ID", "vv_ID"]),
    ("InitialData__InitialData__Type__SphericalGaussian___globals",
     "from nrpytutorial_amd.ScalarWave import InitialData as m\nm.InitialData(Type='SphericalGaussian')\n",
     "m", ["uu_ID", "vv_ID"]),
    ("BSSN_quantities__declare_BSSN_gridfunctions_if_not_declared_already__globals",
     "from nrpytutorial_amd.BSSN import BSSN_quantities as m\nm.declare_BSSN_gridfunctions_if_not_declared_already()\n",
     "m", ['hDD', 'aDD', 'lambdaU', 'vetU', 'betU', 'trK', 'cf', 'alpha']),
    ("BSSN_quantities__BSSN_basic_tensors__globals",
     RFM_INIT + "from nrpytutorial_amd.BSSN import BSSN_quantities as m\nm.BSSN_basic_tensors()\n",
     "m", ['gammabarDD', 'AbarDD', 'LambdabarU', 'betaU', 'BU']),
    ("BSSN_quantities__gammabar__inverse_and_derivs__globals",
     RFM_INIT + "from nrpytutorial_amd.BSSN import BSSN_quantities as m\nm.gammabar__inverse_and_derivs()\n",
     "m", ['gammabarUU', 'gammabarDD_dD', 'gammabarDD_dupD', 'gammabarDD_dDD', 'GammabarUDD']),
    ("BSSN_quantities__detgammabar_and_derivs__globals",
     RFM_INIT + "from nrpytutorial_amd.BSSN import BSSN_quantities as m\nm.detgammabar_and_derivs()\n",
     "m", ['detgammabar', 'detgammabar_dD', 'detgammabar_dDD']),
    ("BSSN_quantities__AbarUU_AbarUD_trAbar_AbarDD_dD__globals",
     RFM_INIT + "from nrpytutorial_amd.BSSN import BSSN_quantities as m\nm.AbarUU_AbarUD_trAbar_AbarDD_dD()\n",
     "m", ['AbarUU', 'AbarUD', 'trAbar', 'AbarDD_dD', 'AbarDD_dupD']),
    ("BSSN_quantities__RicciBar__gammabarDD_dHatD__DGammaUDD__DGammaU__globals",
     RFM_INIT + "from nrpytutorial_amd.BSSN import BSSN_quantities as m\n"
     "m.RicciBar__gammabarDD_dHatD__DGammaUDD__DGammaU()\n",
     "m", ['RbarDD', 'DGammaUDD', 'gammabarDD_dHatD', 'DGammaU']),
    ("BSSN_quantities__betaU_derivs__globals",
     RFM_INIT + "from nrpytutorial_amd.BSSN import BSSN_quantities as m\nm.betaU_derivs()\n",
     "m", ['betaU_dD', 'betaU_dupD', 'betaU_dDD']),
    ("BSSN_quantities__phi_and_derivs__globals",
     RFM_INIT + "from nrpytutorial_amd.BSSN import BSSN_quantities as m\nm.phi_and_derivs()\n",
     "m", ['phi_dD', 'phi_dupD', 'phi_dDD', 'exp_m4phi', 'phi_dBarD', 'phi_dBarDD']),
    ("BSSN_RHSs__BSSN_RHSs__globals",
     "from nrpytutorial_amd.BSSN import BSSN_RHSs as m\nm.BSSN_RHSs()\n",
     "m", ['cf_rhs', 'trK_rhs', 'lambda_rhsU', 'a_rhsDD', 'h_rhsDD']),
    ("gauge_RHSs__BSSN_gauge_RHSs__globals",
     "from nrpytutorial_amd.BSSN import BSSN_RHSs as rhs\nrhs.BSSN_RHSs()\n"
     "from nrpytutorial_amd.BSSN import BSSN_gauge_RHSs as m\nm.BSSN_gauge_RHSs()\n",
     "m", ['alpha_rhs', 'bet_rhsU', 'vet_rhsU']),
    ("BSSN_constraints__BSSN_constraints__globals",
     "from nrpytutorial_amd.BSSN import BSSN_constraints as m\nm.BSSN_constraints()\n",
     "m", ['H', 'MU']),
    ("BSSN_in_terms_of_ADM__gammabarDD_hDD__gammaDD_None__globals",
     RFM_INIT + "from nrpytutorial_amd.BSSN import BSSN_in_terms_of_ADM as m\nm.gammabarDD_hDD(gammaDD=None)\n",
     "m", ['gammabarDD', 'hDD']),
    ("BSSN_in_terms_of_ADM__trK_AbarDD_aDD__gammaDD_None_KDD_None__globals",
     RFM_INIT + "from nrpytutorial_amd.BSSN import BSSN_in_terms_of_ADM as m\nm.trK_AbarDD_aDD(gammaDD=None, KDD=None)\n",
     "m", ['trK', 'AbarDD', 'aDD']),
    ("BSSN_in_terms_of_ADM__LambdabarU_lambdaU__exact_gammaDD__gammaDD_None__globals",
     RFM_INIT + "from nrpytutorial_amd.BSSN import BSSN_in_terms_of_ADM as m\n"
     "m.LambdabarU_lambdaU__exact_gammaDD(gammaDD=None)\n",
     "m", ['LambdabarU', 'lambdaU']),
    ("BSSN_in_terms_of_ADM__cf_from_gammaDD__gammaDD_None__globals",
     RFM_INIT + "from nrpytutorial_amd.BSSN import BSSN_in_terms_of_ADM as m\nm.cf_from_gammaDD(gammaDD=None)\n",
     "m", ['cf']),
    ("BSSN_in_terms_of_ADM__betU_vetU__betaU_None_BU_None__globals",
     RFM_INIT + "from nrpytutorial_amd.BSSN import BSSN_in_terms_of_ADM as m\nm.betU_vetU(betaU=None, BU=None)\n",
     "m", ['vetU', 'betU']),
    ("BrillLindquist__BrillLindquist__ComputeADMGlobalsOnly_True__globals",
     "from nrpytutorial_amd.BSSN import BrillLindquist as m\nm.BrillLindquist(ComputeADMGlobalsOnly=True)\n",
     "m", ['alphaCart', 'betaCartU', 'BCartU', 'gammaCartDD', 'KCartDD']),
]

MAXWELL_SYS = "from nrpytutorial_amd.Maxwell import InitialData\npar.set_parval_from_str('Maxwell.InitialData::System_to_use', '%s')\n"
CASES += [
    ("VacuumMaxwell_Flat_Evol_Cartesian_System_I__VacuumMaxwellRHSs__globals",
     MAXWELL_SYS % "System_I" + "from nrpytutorial_amd.Maxwell import VacuumMaxwell_Flat_Evol_Cartesian as m\n"
     "m.VacuumMaxwellRHSs()\n", "m", ['ArhsU', 'ErhsU', 'C', 'psi_rhs']),
    ("VacuumMaxwell_Flat_Evol_Cartesian_System_II__VacuumMaxwellRHSs__globals",
     MAXWELL_SYS % "System_II" + "from nrpytutorial_amd.Maxwell import VacuumMaxwell_Flat_Evol_Cartesian as m\n"
     "m.VacuumMaxwellRHSs()\n", "m", ['ArhsU', 'ErhsU', 'C', 'psi_rhs', 'Gamma_rhs', 'G']),
    ("VacuumMaxwell_Flat_Evol_Curvilinear_rescaled__VacuumMaxwellRHSs_rescaled__globals",
     MAXWELL_SYS % "System_II" + "from nrpytutorial_amd.Maxwell import VacuumMaxwell_Flat_Evol_Curvilinear_rescaled"
     " as m\nm.VacuumMaxwellRHSs_rescaled()\n", "m",
     ['erhsU', 'arhsU', 'psi_rhs', 'Gamma_rhs', 'C', 'G', 'EU_Cart', 'AU_Cart']),
    ("InitialData_System_II__InitialData__globals",
     MAXWELL_SYS % "System_II" + "m = InitialData\nm.InitialData()\n", "m", ['AidU', 'EidU', 'psi_ID', 'Gamma_ID']),
]

SLOW = {"BSSN_RHSs__BSSN_RHSs__globals", "gauge_RHSs__BSSN_gauge_RHSs__globals",
        "BSSN_constraints__BSSN_constraints__globals"}

DRIVER = r"""
import json, sys
sys.path.insert(0, %(root)r); sys.path.insert(0, %(tests)r)
from nrpytutorial_amd import NRPy_param_funcs as par, reference_metric as rfm, grid as gri
import trusted_eval as te
%(setup)s
glb = {n: getattr(%(mod)s, n) for n in %(names)r}
print("RESULT " + json.dumps(te.compare(%(entry)r, glb)))
"""


def run_case(entry, setup, mod, names):
    code = DRIVER % dict(root=ROOT, tests=HERE, setup=setup, mod=mod, names=names, entry=entry)
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=900)
    assert p.returncode == 0, p.stderr[-3000:]
    line = [l for l in p.stdout.splitlines() if l.startswith("RESULT ")][-1]
    return json.loads(line[len("RESULT "):])


@pytest.mark.parametrize("entry,setup,mod,names", [
    pytest.param(*c, id=c[0].split("__")[0] + ":" + c[0].split("__")[1],
                 marks=[pytest.mark.slow] if c[0] in SLOW else []) for c in CASES])
def test_trusted_values(entry, setup, mod, names):
    failures = run_case(entry, setup, mod, names)
    assert failures == [], "\n".join(failures[:20])


def test_every_fixture_entry_is_covered():
    with open(os.path.join(HERE, "golden", "trusted_values.json")) as f:
        d = json.load(f)
    assert sorted(set(d) - {c[0] for c in CASES}) == []
