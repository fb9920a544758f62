"""Loop generation -> gfx950 launch geometry.

``simple_loop(options, interior)`` keeps the option vocabulary of loop.py
(loop.py:129-194): ``AllPoints`` / ``InteriorPoints`` choose the index range
([0, N+2G) or [G, N+G) in each direction), ``Read_xxs`` makes the coordinates
xx0, xx1, xx2 available, ``Enable_rfm_precompute`` reads the precomputed
reference-metric tables, and ``EnableSIMD`` / ``DisableOpenMP`` /
``OMP_custom_pragma`` are accepted and have no effect.

Instead of a triple loop, the body runs once per lane: i0 (the fastest memory
direction) maps to threadIdx.x so a 64-lane wavefront reads 64 consecutive
doubles of every gridfunction; blocks tile (i0, i1, i2) with a shape set at
launch time (nrpy_set_block_shape) and are remapped so that each XCD works on
a contiguous part of the grid.
``loop1D`` / ``loop`` produce plain C loops for host-side code.
"""
import re
from . import NRPy_param_funcs as par

# Default tile: one wavefront along i0, four i2 planes (measured best for the
# BSSN kernels at 192^3 with the XCD remap: neighbours in i2 come from the same
# block's other waves).
BLOCK_X = 64
BLOCK_Y = 1
BLOCK_Z = 4
# every kernel is compiled for blocks of up to this many threads (the tunable shapes obey it)
MAX_THREADS = 256


# kernel name -> minimum waves per SIMD the compiler must fit its registers for (the
# second __launch_bounds__ argument); set per bundle where a kernel sits just above a
# register-occupancy step
MIN_WAVES_PER_EU = {}


def launch_bounds(name=None):
    """``__launch_bounds__`` argument list. A MIN_WAVES_PER_EU entry for the kernel, or
    NRPY_MIN_WAVES_PER_EU (tuning, every kernel), asks the compiler for at least that many
    waves per SIMD, capping registers per lane."""
    import os
    w = os.environ.get("NRPY_MIN_WAVES_PER_EU", "") or MIN_WAVES_PER_EU.get(name, "")
    return "%d, %d" % (MAX_THREADS, int(w)) if w else "%d" % MAX_THREADS
# kernel name -> id into the runtime's tunable block-shape table (NRPY_MAX_KERNELS entries)
KERNEL_IDS = {}
MAX_KERNEL_IDS = 256
# kernel name -> what is needed to emit its batched (many-instance) variant
BATCH_SPECS = {}


# Nonzero while a bundle generator emits kernels that should read their
# mixed-derivative stencils from LDS (codegen/lds.py); the value is NGHOSTS.
LDS_STAGING = 0
# kernel name -> fixed block shape of every kernel rewritten that way
LDS_KERNELS = {}
# kernel name -> tile shape for its staging, where a generator picks one other than lds.TILE
LDS_TILES = {}


def lds_tile(name):
    from .codegen import lds
    return tuple(LDS_TILES.get(name, lds.TILE))


def stage_lds(name, kernel):
    """(kernel text, launcher block expression or None): when LDS_STAGING is on, the kernel
    with its mixed-derivative stencils staged in LDS boxes and the x-offset reads of its
    other stencil gridfunctions built by cross-lane moves (codegen/lds.py). A kernel with
    neither keeps its tunable block shape; one with only the moves runs on the default
    block shape, fixed. A kernel with a fixed block shape may also get tile-relative
    addressing (codegen/addressing.py)."""
    if not LDS_STAGING:
        return kernel, None
    from .codegen import lds
    tile = lds_tile(name)
    staged = lds.stage_kernel(kernel, LDS_STAGING, tile)
    if staged is None:
        tile = tuple(LDS_TILES.get(name, (BLOCK_X, BLOCK_Y, BLOCK_Z)))
    shifted = None
    if lds.xshift_wanted(name):
        shifted = lds.xshift_kernel(staged if staged is not None else kernel, LDS_STAGING, tile)
    if shifted is not None:
        staged = lds.sink_xshift_values(shifted) if lds.xshift_sink() else shifted
    if staged is None:
        return kernel, None
    LDS_KERNELS[name] = list(tile)
    from .codegen import addressing
    form = addressing.form_for(name)
    if form:
        # LDS_STAGING is the bundle's NGHOSTS (FD order / 2 + 1)
        staged = addressing.tile_address(staged, form, tile, LDS_STAGING, name=name)
    return staged, "dim3(%d, %d, %d)" % tile


def kernel_id(name):
    if name not in KERNEL_IDS:
        if len(KERNEL_IDS) >= MAX_KERNEL_IDS:
            raise ValueError("more than %d tunable kernels in one bundle: raise NRPY_MAX_KERNELS "
                             "(csrc/nrpy_runtime_pre.h) and MAX_KERNEL_IDS together" % MAX_KERNEL_IDS)
        KERNEL_IDS[name] = len(KERNEL_IDS)
    return KERNEL_IDS[name]


def loop1D(idx_var="i", lower_bound="0", upper_bound="N", increment="1", pragma="#pragma omp parallel for"):
    header = "%sfor (int %s = %s; %s < %s; %s += %s) {" % (
        (pragma + "\n") if pragma else "", idx_var, lower_bound, idx_var, upper_bound, idx_var, increment)
    return header, "} // END LOOP: for (int %s = %s; %s < %s; %s += %s)" % (
        idx_var, lower_bound, idx_var, upper_bound, idx_var, increment)


def loop(idx_var, lower_bound, upper_bound, increment, pragma, padding="  ", interior="", tile_size=""):
    if not isinstance(idx_var, list):
        idx_var, lower_bound, upper_bound, increment, pragma = \
            [idx_var], [lower_bound], [upper_bound], [increment], [pragma]
    heads, tails = [], []
    for i in range(len(idx_var)):
        h, t = loop1D(idx_var[i], lower_bound[i], upper_bound[i], increment[i], pragma[i])
        heads.append(padding * i + h)
        tails.append(padding * i + t)
    body = "\n".join(padding * len(idx_var) + l for l in interior.splitlines())
    return "\n".join(heads) + ("\n" + body if body else "") + "\n" + "\n".join(reversed(tails)) + "\n"


def parse_loopopts(options):
    opts = {"range": None, "Read_xxs": False, "rfm": False}
    for o in [s.strip() for s in options.split(",") if s.strip()]:
        if o == "AllPoints":
            opts["range"] = "all"
        elif o == "InteriorPoints":
            opts["range"] = "interior"
        elif o == "Read_xxs":
            opts["Read_xxs"] = True
        elif o == "Enable_rfm_precompute":
            opts["rfm"] = True
        elif o in ("EnableSIMD", "DisableOpenMP", "oldloops") or o.startswith("OMP_custom_pragma") \
                or o.startswith("OMP_collapse"):
            pass
        else:
            raise ValueError("unknown loop option '" + o + "'")
    if opts["range"] is None:
        raise ValueError("loop options must contain AllPoints or InteriorPoints")
    if opts["Read_xxs"] and opts["rfm"]:
        raise ValueError("Read_xxs and Enable_rfm_precompute are mutually exclusive")
    return opts


def rfm_reads(body):
    """Statements reading precomputed reference-metric tables used in ``body``."""
    from . import reference_metric as rfm
    out = []
    for name, expr, axes in rfm.rfm_precompute_tables():
        if not re.search(r"\b%s\b" % re.escape(name), body):
            continue
        if len(axes) == 1:
            out.append("const double %s = rfmstruct.%s[i%d];" % (name, name, axes[0]))
        elif axes == (0, 1):
            out.append("const double %s = rfmstruct.%s[i0 + Nxx_plus_2NGHOSTS0*i1];" % (name, name))
        else:
            raise ValueError("unsupported reference-metric table dependence for " + name)
    return out


def simple_loop(options, interior, batched=False):
    """Kernel body: lane -> gridpoint mapping, bounds check, coordinate/table reads, ``interior``.

    batched=True: the block's tile comes from the instance table of a batched launch
    (NRPY_BATCH_TILE, defined by the batched kernel prologue) instead of the XCD remap."""
    o = parse_loopopts(options)
    lo = "NGHOSTS" if o["range"] == "interior" else "0"
    n = ("Nxx%d" if o["range"] == "interior" else "Nxx_plus_2NGHOSTS%d")
    # One-dimensional grid of blocks; NRPY_BLOCK_COORDS remaps the hardware's
    # round-robin block-to-XCD assignment to contiguous tiles per XCD and
    # decodes the (bx, by, bz) tile of this block.
    if batched:
        s = "NRPY_BATCH_TILE(%s, %s, %s);\n" % (n % 0, n % 1, n % 2)
    else:
        s = "NRPY_BLOCK_COORDS(%s, %s, %s);\n" % (n % 0, n % 1, n % 2)
    for d, t in ((0, "x"), (1, "y"), (2, "z")):
        s += "const int i%d = %s + nrpy_b%s*(int)blockDim.%s + (int)threadIdx.%s;\n" % (d, lo, t, t, t)
    if o["range"] == "interior":
        s += "if (i0 >= Nxx0 + NGHOSTS || i1 >= Nxx1 + NGHOSTS || i2 >= Nxx2 + NGHOSTS) return;\n"
    else:
        s += ("if (i0 >= Nxx_plus_2NGHOSTS0 || i1 >= Nxx_plus_2NGHOSTS1 || i2 >= Nxx_plus_2NGHOSTS2) return;\n")
    if o["Read_xxs"]:
        for d in range(3):
            if re.search(r"\bxx%d\b" % d, interior):
                s += "const double xx%d = xx%d_arr[i%d];\n" % (d, d, d)
    if o["rfm"]:
        s += "\n".join(rfm_reads(interior)) + "\n"
    return s + interior


def _used_cparams(text):
    names = set(re.findall(r"\b[A-Za-z_][A-Za-z_0-9]*\b", text))
    out = []
    for name, ctype in par.paramstruct_members():
        if name in ("xx0", "xx1", "xx2"):
            continue
        if name in names:
            out.append((name, ctype))
    return out


def _ctype(c):
    return {"REAL": "double", "int": "int", "bool": "int", "char": "const char *"}.get(c, c)


def kernel_and_launcher(name, desc, args, preloop, body, loopopts, postloop):
    """Assemble a ``__global__`` kernel and its ``extern "C"`` launcher.

    ``args`` is the reference-style parameter list [(ctype, name)]:
    ``paramstruct *params`` and ``rfm_struct *rfmstruct`` become by-value
    kernel arguments, ``REAL *xx[3]`` becomes three coordinate arrays and every
    other pointer is a device array. The launcher appends ``hipStream_t stream``
    and returns 0 on success or a negative error code."""
    o = parse_loopopts(loopopts)
    kbody = simple_loop(loopopts, body)
    full = preloop + "\n" + kbody + "\n" + postloop
    arrays, has_rfm, batchable = [], False, True
    for ctype, an in args:
        base = ctype.replace("const", "").replace("*", "").strip()
        if base == "paramstruct":
            continue
        if base == "rfm_struct":
            has_rfm = True
        elif an == "xx" and "[3]" in ctype:
            batchable = False
        else:
            arrays.append(("const double *__restrict__ " if "const" in ctype else "double *__restrict__ ", an))
    if batchable:
        BATCH_SPECS[name] = {"arrays": arrays, "rfm": has_rfm, "template": None,
                             "body": preloop + "\n" + body + "\n" + postloop if (preloop or postloop) else body,
                             "loopopts": loopopts, "dt": False}
    kargs, largs, call = [], [], []
    for ctype, an in args:
        base = ctype.replace("const", "").replace("*", "").strip()
        if base == "paramstruct":
            kargs.append("const paramstruct params")
            largs.append("const paramstruct *params")
            call.append("*params")
        elif base == "rfm_struct":
            kargs.append("const rfm_struct rfmstruct")
            largs.append("const rfm_struct *rfmstruct")
            call.append("*rfmstruct")
        elif an == "xx" and "[3]" in ctype:
            kargs += ["const double *__restrict__ xx%d_arr" % d for d in range(3)]
            largs.append("double *xx[3]")
            call += ["xx[%d]" % d for d in range(3)]
        else:
            isconst = "const" in ctype
            kargs.append(("const double *__restrict__ " if isconst else "double *__restrict__ ") + an)
            largs.append(("const double *" if isconst else "double *") + an)
            call.append(an)
    needs = _used_cparams(full)
    unpack = "".join("  const %s %s = params.%s;\n" % (_ctype(t), n, n) for n, t in needs)
    # always needed for bounds and indexing
    for n in ("Nxx0", "Nxx1", "Nxx2", "Nxx_plus_2NGHOSTS0", "Nxx_plus_2NGHOSTS1", "Nxx_plus_2NGHOSTS2"):
        if not any(n == x for x, _ in needs):
            unpack += "  const int %s = params.%s;\n" % (n, n)
    kname = name + "__gpu"
    kernel = ("/* %s */\n__global__ void __launch_bounds__(%s) %s(%s) {\n%s%s}\n" %
              (desc, launch_bounds(name), kname, ", ".join(kargs), unpack,
               "".join("  " + l + "\n" for l in full.splitlines())))
    rng = "interior" if o["range"] == "interior" else "all"
    block = None
    if rng == "interior":
        kernel, block = stage_lds(name, kernel)
    if block is None:
        from .codegen import addressing
        form = addressing.form_for(name)
        if form:
            # tile-relative addressing needs a fixed tile: the default block shape
            tile = (BLOCK_X, BLOCK_Y, BLOCK_Z)
            kernel = addressing.tile_address(kernel, form, tile, name=name)
            block = "dim3(%d, %d, %d)" % tile
    if rng == "interior":
        n0, n1, n2 = "params->Nxx0", "params->Nxx1", "params->Nxx2"
    else:
        n0, n1, n2 = "params->Nxx_plus_2NGHOSTS0", "params->Nxx_plus_2NGHOSTS1", "params->Nxx_plus_2NGHOSTS2"
    ptr_checks = " || ".join("!" + c for c in call if not c.startswith("*") and "xx[" not in c) or "0"
    kid = kernel_id(name)
    launcher = (
        'extern "C" int %s(%s, hipStream_t stream) {\n'
        "  if (!params) return NRPY_ERR_NULL;\n"
        "  if (%s) return NRPY_ERR_NULL;\n"
        "  if (nrpy_check_params(params)) return NRPY_ERR_SHAPE;\n"
        "  const dim3 block = %s;\n"
        "  const dim3 grid(nrpy_nblocks(block, %s, %s, %s));\n"
        "  hipLaunchKernelGGL(%s, grid, block, 0, stream, %s);\n"
        "  return hipGetLastError() == hipSuccess ? 0 : NRPY_ERR_LAUNCH;\n"
        "}\n" % (name, ", ".join(largs), ptr_checks,
                 block or "nrpy_block(%d, %d, %d, %d)" % (kid, BLOCK_X, BLOCK_Y, BLOCK_Z), n0, n1, n2, kname,
                 ", ".join(call)))
    proto = 'extern "C" int %s(%s, hipStream_t stream);' % (name, ", ".join(largs))
    return kernel, launcher, proto
