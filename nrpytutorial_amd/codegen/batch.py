"""Batched variants of generated kernels: one launch advances many instances.

A batched kernel has the same per-point body as its single-grid kernel. Its
prologue differs: the block looks up its instance in the launch's tile prefix
table (NRPY_BATCH_FIND), takes that instance's paramstruct, rfm tables, time
step and arrays from the device instance table, and decodes its tile inside
the instance's index box (NRPY_BATCH_TILE). Arrays are passed as roles (y_n,
k_odd, k_even, running total, auxevol, aux), so one compiled kernel serves
every substep of the RK4 schedule.
"""
import re

from .. import loop as lp

ROLES = ("NRPY_BATCH_ROLE_YN", "NRPY_BATCH_ROLE_KODD", "NRPY_BATCH_ROLE_KEVEN", "NRPY_BATCH_ROLE_RT",
         "NRPY_BATCH_ROLE_AUXEVOL", "NRPY_BATCH_ROLE_AUX")


def instance_struct():
    return ("#define NRPY_BATCH_ROLES 6\n"
            "typedef struct {\n  paramstruct params;\n  rfm_struct rfm;\n  double *arr[NRPY_BATCH_ROLES];\n"
            "  double dt;\n  long Ntot;\n} nrpy_batch_instance;\n")


def batched_kernel(name, kid):
    """(kernel source, launcher source, range) for the batched variant of kernel ``name``."""
    spec = lp.BATCH_SPECS[name]
    o = lp.parse_loopopts(spec["loopopts"])
    rng = "NRPY_RANGE_INTERIOR" if o["range"] == "interior" else "NRPY_RANGE_ALL"
    kbody = lp.simple_loop(spec["loopopts"], spec["body"], batched=True)
    unpack = "".join("  const %s %s = params.%s;\n" % (lp._ctype(t), n, n) for n, t in lp._used_cparams(kbody))
    for n in ("Nxx0", "Nxx1", "Nxx2", "Nxx_plus_2NGHOSTS0", "Nxx_plus_2NGHOSTS1", "Nxx_plus_2NGHOSTS2"):
        if not re.search(r"const int %s = " % n, unpack):
            unpack += "  const int %s = params.%s;\n" % (n, n)
    kargs = ["const nrpy_batch_instance *__restrict__ batch", "const int *__restrict__ tile_prefix", "const int ninst"]
    kargs += ["const int role_%s" % an for _, an in spec["arrays"]]
    pro = "  NRPY_BATCH_FIND();\n  const paramstruct params = batch[nrpy_k].params;\n"
    if spec["rfm"]:
        pro += "  const rfm_struct rfmstruct = batch[nrpy_k].rfm;\n"
    if spec["dt"]:
        pro += "  const double dt = batch[nrpy_k].dt;\n"
    for q, an in spec["arrays"]:
        pro += "  %s%s = batch[nrpy_k].arr[role_%s];\n" % (q, an, an)
    tmpl = "template <int S>\n" if spec["template"] else ""
    kname = name + "__gpu_batched"
    kernel = ("%s__global__ void __launch_bounds__(%s) %s(%s) {\n%s%s%s}\n"
              % (tmpl, lp.launch_bounds(name), kname, ", ".join(kargs), pro, unpack,
                 "".join("  " + l + "\n" for l in kbody.splitlines())))
    roles = ", ".join("const int role_%s" % an for _, an in spec["arrays"])
    role_call = ", ".join("role_%s" % an for _, an in spec["arrays"])
    checks = " || ".join("role_%s < 0 || role_%s >= NRPY_BATCH_ROLES" % (an, an) for _, an in spec["arrays"]) or "0"
    if spec["template"]:
        cases = "".join("    case %d: hipLaunchKernelGGL(%s<%d>, dim3(nb), block, 0, stream, b->d_inst, prefix, "
                        "b->ninst, %s); break;\n" % (s, kname, s, role_call) for s in range(4))
        launch = "  switch (S) {\n%s    default: return NRPY_ERR_PARAM;\n  }\n" % cases
        sig = "const int S, nrpy_batch *b, %s, hipStream_t stream" % roles
    else:
        launch = ("  hipLaunchKernelGGL(%s, dim3(nb), block, 0, stream, b->d_inst, prefix, b->ninst, %s);\n"
                  % (kname, role_call))
        sig = "nrpy_batch *b, %s, hipStream_t stream" % roles
    launcher = (
        'extern "C" int %s_batched(%s) {\n'
        "  if (!b) return NRPY_ERR_NULL;\n"
        "  if (%s) return NRPY_ERR_PARAM;\n"
        "  const dim3 block = nrpy_block(%d, %d, %d, %d);\n"
        "  unsigned nb = 0;\n"
        "  const int *prefix = nrpy_batch_prefix(b, block, %s, &nb);\n"
        "  if (!prefix) return NRPY_ERR_PARAM;  // block shape changed since nrpy_batch_prepare\n"
        "%s"
        "  return hipGetLastError() == hipSuccess ? NRPY_OK : NRPY_ERR_LAUNCH;\n"
        "}\n" % (name, sig, checks, kid, BATCH_BLOCK[0], BATCH_BLOCK[1], BATCH_BLOCK[2], rng, launch))
    return kernel, launcher, rng


# Default tile of batched launches: the instances are small (tens of points per
# direction, Nphi = 2), so the tile spans i0 and i1 only.
BATCH_BLOCK = (64, 4, 1)


def batched_kernels(names):
    """Source for the batched variants of ``names`` and the kernel table nrpy_batch.h needs."""
    src, table = "", []
    for n in names:
        kid = lp.kernel_id(n + "_batched")
        k, l, rng = batched_kernel(n, kid)
        src += k + l
        table.append("{%d, %s, %d, %d, %d}" % (kid, rng, BATCH_BLOCK[0], BATCH_BLOCK[1], BATCH_BLOCK[2]))
    table_src = "static const int nrpy_batched_kernels[][5] = {%s};\n" % ", ".join(table)
    return src, table_src
