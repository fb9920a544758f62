"""Write tests/golden/trusted_values.json from the reference's trusted-value dicts.

The reference stores known answers for its symbolic expressions in files such
as ScalarWave/tests/trusted_values_dict.py, one assignment per entry:
``trusted_values_dict['<entry>'] = {'<global>': mpf('<digits>'), ...}``.
This script reads those files as text with ``ast`` (nothing is imported or
executed) and keeps the entries on the FD-RHS hot path. Run it where the
reference checkout is present:

    python tests/golden/extract_trusted_values.py /root/reference
"""
import ast
import json
import os
import sys

SOURCES = {
    "ScalarWave": "ScalarWave/tests/trusted_values_dict.py",
    "Maxwell": "Maxwell/tests/trusted_values_dict.py",
    "BSSN": "BSSN/tests/trusted_values_dict.py",
    "rfm": "tests/trusted_values_dict.py",
}

KEEP = {
    "ScalarWave": ["ScalarWave_RHSs__ScalarWave_RHSs__globals",
                   "ScalarWaveCurvilinear_RHSs__ScalarWaveCurvilinear_RHSs__globals",
                   "InitialData__InitialData__Type__PlaneWave___globals",
                   "InitialData__InitialData__Type__SphericalGaussian___globals"],
    "Maxwell": ["VacuumMaxwell_Flat_Evol_Cartesian_System_I__VacuumMaxwellRHSs__globals",
                "VacuumMaxwell_Flat_Evol_Cartesian_System_II__VacuumMaxwellRHSs__globals",
                "VacuumMaxwell_Flat_Evol_Curvilinear_rescaled__VacuumMaxwellRHSs_rescaled__globals",
                "InitialData_System_II__InitialData__globals"],
    "BSSN": ["BSSN_RHSs__BSSN_RHSs__globals",
             "gauge_RHSs__BSSN_gauge_RHSs__globals",
             "BSSN_constraints__BSSN_constraints__globals",
             "BrillLindquist__BrillLindquist__ComputeADMGlobalsOnly_True__globals",
             "BSSN_quantities__declare_BSSN_gridfunctions_if_not_declared_already__globals",
             "BSSN_quantities__BSSN_basic_tensors__globals",
             "BSSN_quantities__gammabar__inverse_and_derivs__globals",
             "BSSN_quantities__detgammabar_and_derivs__globals",
             "BSSN_quantities__AbarUU_AbarUD_trAbar_AbarDD_dD__globals",
             "BSSN_quantities__RicciBar__gammabarDD_dHatD__DGammaUDD__DGammaU__globals",
             "BSSN_quantities__betaU_derivs__globals",
             "BSSN_quantities__phi_and_derivs__globals",
             "BSSN_in_terms_of_ADM__gammabarDD_hDD__gammaDD_None__globals",
             "BSSN_in_terms_of_ADM__trK_AbarDD_aDD__gammaDD_None_KDD_None__globals",
             "BSSN_in_terms_of_ADM__LambdabarU_lambdaU__exact_gammaDD__gammaDD_None__globals",
             "BSSN_in_terms_of_ADM__cf_from_gammaDD__gammaDD_None__globals",
             "BSSN_in_terms_of_ADM__betU_vetU__betaU_None_BU_None__globals"],
    "rfm": ["rfm_Cartesian__reference_metric__True__globals",
            "rfm_Spherical__reference_metric__True__globals",
            "rfm_SinhSpherical__reference_metric__True__globals",
            "rfm_SinhSphericalv2__reference_metric__True__globals",
            "rfm_Cylindrical__reference_metric__True__globals",
            "rfm_SinhCylindrical__reference_metric__True__globals",
            "rfm_SinhCylindricalv2__reference_metric__True__globals",
            "rfm_SymTP__reference_metric__True__globals",
            "rfm_SinhSymTP__reference_metric__True__globals"],
}


def _value(node):
    """mpf('x') -> 'x';  mpc(real='a', imag='b') -> ['a', 'b']."""
    if not isinstance(node, ast.Call) or not isinstance(node.func, ast.Name):
        raise ValueError("unexpected value node")
    if node.func.id == "mpf":
        return node.args[0].value
    if node.func.id == "mpc":
        kw = {k.arg: k.value.value for k in node.keywords}
        return [kw["real"], kw["imag"]]
    raise ValueError("unexpected call " + node.func.id)


def parse(path):
    tree = ast.parse(open(path).read())
    out = {}
    for stmt in tree.body:
        if not isinstance(stmt, ast.Assign) or not isinstance(stmt.targets[0], ast.Subscript):
            continue
        key = stmt.targets[0].slice.value
        if not isinstance(stmt.value, ast.Dict):
            continue
        out[key] = {k.value: _value(v) for k, v in zip(stmt.value.keys, stmt.value.values)}
    return out


def main(refroot):
    result = {}
    for group, rel in SOURCES.items():
        entries = parse(os.path.join(refroot, rel))
        for key in KEEP[group]:
            result[key] = {"source": rel, "values": entries[key]}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "trusted_values.json")
    with open(dst, "w") as f:
        json.dump(result, f, indent=0, sort_keys=True)
    print("wrote", dst, len(result), "entries")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
