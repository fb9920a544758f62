"""Stage the mixed-derivative stencils of an interior-point kernel in LDS.

The BSSN kernels read each input gridfunction at up to 67 stencil points
(mixed second derivatives are 16-point stencils per coordinate plane). At the
register pressure of these kernels (one or two waves per SIMD) the compiler
issues such a stencil's loads one at a time, each followed by a wait for the
memory round trip. This pass moves the gridfunctions that carry mixed
derivatives into LDS once per tile, before any other work, so those stencil
reads become LDS reads with immediate offsets (a fraction of the L2 latency)
and no address arithmetic.

Layout. A block is a fixed tile of interior points, one lane per point:
TILE = 16x4x4 unless the generator picks another shape for a kernel
(loop.LDS_TILES). For every staged gridfunction the block holds the box
(TX + 2*NGHOSTS) x (TY + 2H) x (TZ + 2H) around the tile in LDS, x fastest,
where H is the largest stencil offset of the staged reads. The box starts
NGHOSTS points before the tile in x, so its rows begin at an even index
(TX is even).

Loading. When every box row is 16-byte aligned in global memory (even row
length Nxx_plus_2NGHOSTS0 and a 16-byte aligned array), each wave copies
64 x 16 bytes per instruction straight into LDS (``global_load_lds_dwordx4``:
lane l writes bytes [16 l, 16 l + 16) after the wave's LDS base), so the
copy uses no registers. Otherwise the block copies element by element
through registers. Box points outside the array are not loaded; no active
point reads them. One barrier follows the copy, ahead of the kernel's
bounds check, so every thread of the block reaches it.

The pass rewrites generated kernel text: the reads it recognises are the
``const double v = in_gfs[IDX4(GF, i0+a, i1+b, i2+c)];`` statements the
finite-difference emitter writes. Arithmetic is untouched, so results are
bit-identical to the unstaged kernel.
"""
import re

TILE = (16, 4, 4)
# LDS bytes a block may use: 160 KiB per CU, and kernels at two waves per SIMD
# run two 256-thread blocks per CU.
MAX_LDS_BYTES = 78 * 1024

_READ = re.compile(r"^(\s*const double \w+ = )(in_gfs|auxevol_gfs)\[IDX4\((\w+),\s*i0([+-]\d+)?,\s*i1([+-]\d+)?,"
                   r"\s*i2([+-]\d+)?\)\];$")
_BOUNDS = re.compile(r"^\s*if \(i0 >= Nxx0 \+ NGHOSTS \|\| i1 >= Nxx1 \+ NGHOSTS \|\| i2 >= Nxx2 \+ NGHOSTS\) return;$")
# The same reads and bounds check in the Einstein Toolkit form of a kernel: one array per
# gridfunction, ``hDD00GF[CCTK_GFINDEX3D(cctkGH, i0-1,i1,i2)]``, interior [nrpy_lo, nrpy_hi).
_READ_ETK = re.compile(r"^(\s*const double \w+ = )(\w+GF)\[CCTK_GFINDEX3D\(cctkGH,\s*i0([+-]\d+)?,\s*i1([+-]\d+)?,"
                       r"\s*i2([+-]\d+)?\)\];$")
_BOUNDS_ETK = re.compile(r"^\s*if \(i0 >= nrpy_hi0 \|\| i1 >= nrpy_hi1 \|\| i2 >= nrpy_hi2\) return;$")


def _match(line):
    """(statement prefix, array, gridfunction, offsets) of a stencil read, or None. In the
    Einstein Toolkit form the array is None and the gridfunction is its pointer's name."""
    m = _READ.match(line)
    if m:
        return m.group(1), m.group(2), m.group(3), tuple(int(x) if x else 0 for x in m.group(4, 5, 6))
    m = _READ_ETK.match(line)
    if m:
        return m.group(1), None, m.group(2), tuple(int(x) if x else 0 for x in m.group(3, 4, 5))
    return None


def _reads(lines):
    """{gf: (array, set of offsets)} and first-appearance order of the stencil reads."""
    reads, order = {}, []
    for line in lines:
        m = _match(line)
        if m:
            gf = m[2]
            if gf not in reads:
                reads[gf] = (m[1], set())
                order.append(gf)
            reads[gf][1].add(m[3])
    return reads, order


def select(lines, nghosts, tile=TILE):
    """Gridfunctions to stage (those read at a point off two axes at once, i.e. with a
    mixed derivative), their y/z halo H and LDS box dimensions; None if nothing qualifies
    or the boxes would not fit."""
    reads, order = _reads(lines)
    staged = [g for g in order if any(sum(1 for x in d if x) >= 2 for d in reads[g][1])]
    if not staged:
        return None
    H = max(max(abs(x) for x in d) for g in staged for d in reads[g][1])
    box = (tile[0] + 2 * nghosts, tile[1] + 2 * H, tile[2] + 2 * H)
    while staged and len(staged) * box[0] * box[1] * box[2] * 8 > MAX_LDS_BYTES:
        staged.pop()
    if not staged:
        return None
    return [(g, reads[g][0]) for g in staged], H, box


# Index-space names of the two kernel forms: where the interior starts, the extent of
# the allocated box (bounds), the row and plane strides, and when the 16-byte copy
# applies: box rows start at an even index, and the row length is even so that a
# pair (x, x + 1) with x inside the row lies inside it entirely (the copy skips a
# pair that crosses the row's end, so an odd row length would lose its last point).
_SENR = dict(lo=("NGHOSTS", "NGHOSTS", "NGHOSTS"),
             n=("Nxx_plus_2NGHOSTS0", "Nxx_plus_2NGHOSTS1", "Nxx_plus_2NGHOSTS2"),
             stride=("Nxx_plus_2NGHOSTS0", "Nxx_plus_2NGHOSTS1"),
             even="(Nxx_plus_2NGHOSTS0 & 1) == 0")
_ETK = dict(lo=("nrpy_lo0", "nrpy_lo1", "nrpy_lo2"),
            n=("nrpy_lsh0", "nrpy_lsh1", "nrpy_lsh2"),
            stride=("cctkGH.ash0", "cctkGH.ash1"),
            even="((cctkGH.ash0 | nrpy_lsh0 | (nrpy_lo0 - NGHOSTS)) & 1) == 0")


def _prologue(staged, H, box, nghosts, etk=False, tile=TILE):
    TX, TY, TZ = tile
    BX, BY, BZ = box
    nbox = BX * BY * BZ
    nchunk = nbox // 2            # 16-byte chunks (BX is even)
    threads = TX * TY * TZ
    f = _ETK if etk else _SENR
    lo, n, st = f["lo"], f["n"], f["stride"]
    arrays = sorted(set(g if etk else a for g, a in staged))
    aligned = " && ".join("(((size_t)%s & 15) == 0)" % a for a in arrays)

    def src(g, a):
        return g if etk else "%s + (size_t)%s * ntot_" % (a, g)
    s = []
    s.append("__shared__ __attribute__((aligned(16))) double nrpy_lds[%d][%d];" % (len(staged), nbox))
    s.append("{")
    s.append("  /* copy the (%d x %d x %d)-point box around this %dx%dx%d tile of %s to LDS */"
             % (BX, BY, BZ, TX, TY, TZ, ", ".join(g[:-2] if g.endswith("GF") else g for g, _ in staged)))
    s.append("  const int t_ = (int)threadIdx.x + %d * ((int)threadIdx.y + %d * (int)threadIdx.z);" % (TX, TY))
    x0 = "%s - NGHOSTS + nrpy_bx * %d" % (lo[0], TX) if etk else "nrpy_bx * %d" % TX
    s.append("  const int x0_ = %s, y0_ = %s - %d + nrpy_by * %d, z0_ = %s - %d + nrpy_bz * %d;"
             % (x0, lo[1], H, TY, lo[2], H, TZ))
    if not etk:
        s.append("  const size_t ntot_ = (size_t)Nxx_plus_2NGHOSTS0 * Nxx_plus_2NGHOSTS1 * Nxx_plus_2NGHOSTS2;")
    s.append("  if (%s && %s) {" % (f["even"], aligned))
    s.append("    const int w_ = t_ >> 6, l_ = t_ & 63;")
    s.append("    for (int it_ = 0; it_ < %d; it_++) {" % ((nchunk + threads - 1) // threads))
    s.append("      const int c_ = (it_ * %d + w_) * 64 + l_;" % (threads // 64))
    s.append("      const int row_ = c_ / %d, xc_ = c_ - row_ * %d;" % (BX // 2, BX // 2))
    s.append("      const int z_ = row_ / %d, y_ = row_ - z_ * %d;" % (BY, BY))
    s.append("      if (c_ < %d && z0_ + z_ < %s && y0_ + y_ < %s && x0_ + 2 * xc_ + 1 < %s) {"
             % (nchunk, n[2], n[1], n[0]))
    s.append("        const size_t off_ = ((size_t)(z0_ + z_) * %s + (y0_ + y_)) * %s + x0_ + 2 * xc_;" % (st[1], st[0]))
    for q, (g, a) in enumerate(staged):
        s.append("        __builtin_amdgcn_global_load_lds((const void *)(%s + off_), "
                 "(__attribute__((address_space(3))) void *)&nrpy_lds[%d][(it_ * %d + w_) * 128], 16, 0, 0);"
                 % (src(g, a), q, threads // 64))
    s.append("      }")
    s.append("    }")
    s.append("  } else {")
    s.append("    for (int e_ = t_; e_ < %d; e_ += %d) {" % (nbox, threads))
    s.append("      const int row_ = e_ / %d, x_ = e_ - row_ * %d;" % (BX, BX))
    s.append("      const int z_ = row_ / %d, y_ = row_ - z_ * %d;" % (BY, BY))
    s.append("      if (z0_ + z_ < %s && y0_ + y_ < %s && x0_ + x_ < %s) {" % (n[2], n[1], n[0]))
    s.append("        const size_t off_ = ((size_t)(z0_ + z_) * %s + (y0_ + y_)) * %s + x0_ + x_;" % (st[1], st[0]))
    for q, (g, a) in enumerate(staged):
        s.append("        nrpy_lds[%d][e_] = %s;" % (q, "%s[off_]" % g if etk else "%s[(size_t)%s * ntot_ + off_]" % (a, g)))
    s.append("      }")
    s.append("    }")
    s.append("  }")
    s.append("}")
    s.append("__syncthreads();")
    s.append("const int nrpy_li = (((int)threadIdx.z + %d) * %d + ((int)threadIdx.y + %d)) * %d + (int)threadIdx.x + %d;"
             % (H, BY, H, BX, nghosts))
    return s


def stage_kernel(kernel_src, nghosts, tile=TILE):
    """Kernel text with its mixed-derivative stencils staged in LDS, or None when the
    kernel has none (or no recognisable bounds check). The result must be launched
    with blockDim = tile (TILE unless the caller picks another shape)."""
    lines = kernel_src.split("\n")
    sel = select(lines, nghosts, tile)
    if sel is None:
        return None
    staged, H, box = sel
    etk = staged[0][1] is None
    if any((a is None) != etk for _, a in staged):
        return None
    bounds = [k for k, line in enumerate(lines) if (_BOUNDS_ETK if etk else _BOUNDS).match(line)]
    if len(bounds) != 1:
        return None
    q_of = {g: q for q, (g, _) in enumerate(staged)}
    BX, BY, _ = box
    indent = re.match(r"^(\s*)", lines[bounds[0]]).group(1)
    out = []
    for k, line in enumerate(lines):
        if k == bounds[0]:
            out += [indent + p for p in _prologue(staged, H, box, nghosts, etk, tile)]
            out.append(line)
            continue
        m = _match(line)
        if m and m[2] in q_of:
            dx, dy, dz = m[3]
            out.append("%snrpy_lds[%d][nrpy_li + %d];" % (m[0], q_of[m[2]], (dz * BY + dy) * BX + dx))
            continue
        out.append(line)
    return "\n".join(out)


# ---------------------------------------------------------------------------
# x-neighbours by cross-lane moves
# ---------------------------------------------------------------------------
# Every stencil read of a gridfunction that is not staged as a box is a global load,
# and these kernels spend most of their time in the texture addresser (tens of cycles
# per 64-lane load). An x-offset read (i0 +- 1..3 at the same i1, i2) loads a value
# that a neighbouring lane of the same row holds as its centre value. This pass keeps
# two loads per such gridfunction, the centre value C and one halo value H (the first
# HX lanes of a row load i0 - HX, the others i0 + HX), and builds the x-offset values
# with DPP lane moves (nrpy_xshift in nrpy_runtime_pre.h): f(i0 + k) is C of lane
# l + k of the row, and for the lanes at the row's ends whose neighbour lies outside
# it, H of the lane that loaded exactly that point. A row is TX = 16 lanes
# (row_shl/row_shr) or the whole 64-lane wave (wave_shl/wave_shr by one lane, chained).
#
# A lane move reads nothing from a lane that has exited, so every row must have all
# its lanes, or at least 2 HX of them when the grid is narrower than one row. The last
# tile along x is therefore moved left to end at the grid's last point and overlaps
# the tile before it; its lanes over points of that tile compute them again and send
# their stores to a discard array (a pointer select, NRPY_OWNED: a branch around the
# stores would let the compiler move the whole computation behind it, while the lane
# moves, which cannot enter divergent code, stay in front). The in-place RK4 update
# (the output aliases y_n in the last substep) stays race-free: every point is stored
# by exactly one lane. Values are unchanged, so results are bit-identical; the bundle
# rejects grids with fewer than MIN_NXX0 points along x.

# kernels (names) whose x-neighbours are built by lane moves; NRPY_XSHIFT (comma list or
# "all") and NRPY_XSHIFT_SINK=0 override for measurements
XSHIFT_KERNELS = set()
MIN_NXX0 = 0          # set when a kernel is rewritten: 2 HX


def xshift_wanted(name):
    import os
    env = os.environ.get("NRPY_XSHIFT")
    if env is not None:
        return env == "all" or name in env.split(",")
    return name in XSHIFT_KERNELS


def xshift_sink():
    import os
    return os.environ.get("NRPY_XSHIFT_SINK", "1") != "0"
_STORE = re.compile(r"^(\s*)(\w+)\[(IDX4\(\w+, i0, i1, i2\))\] = (.*);$")
_I0 = re.compile(r"^(\s*)const int i0 = NGHOSTS \+ nrpy_bx\*\(int\)blockDim\.x \+ \(int\)threadIdx\.x;$")


def xshift_kernel(kernel_src, nghosts, tile):
    """Kernel text with the x-offset stencil reads of its unstaged gridfunctions built from
    the centre and one halo load by cross-lane moves, or None when nothing qualifies.
    The result must be launched with blockDim = tile."""
    global MIN_NXX0
    TX = tile[0]
    if TX not in (16, 64) or (tile[0] * tile[1] * tile[2]) % 64:
        return None
    lines = kernel_src.split("\n")
    bounds = [k for k, line in enumerate(lines) if _BOUNDS.match(line)]
    i0line = [k for k, line in enumerate(lines) if _I0.match(line)]
    if len(bounds) != 1 or len(i0line) != 1:
        return None
    reads, order = {}, []
    for line in lines:
        m = _READ.match(line)
        if not m:
            continue
        gf = m.group(3)
        if gf not in reads:
            reads[gf] = (m.group(2), set())
            order.append(gf)
        reads[gf][1].add(tuple(int(x) if x else 0 for x in m.group(4, 5, 6)))
    hx = {}
    for g in order:
        xs = [abs(d[0]) for d in reads[g][1] if d[0] and not d[1] and not d[2]]
        if len(xs) >= 2:
            hx[g] = max(xs)
    if not hx or max(hx.values()) > nghosts:
        return None
    MIN_NXX0 = max(MIN_NXX0, 2 * max(hx.values()))
    q_of = {g: q for q, g in enumerate(g for g in order if g in hx)}
    out, done, moved = [], set(), []
    for k, line in enumerate(lines):
        if k == i0line[0]:
            ind = _I0.match(line).group(1)
            out += [ind + "const int nrpy_xt = nrpy_bx * %d;  /* first x of this tile's own points */" % TX,
                    ind + "const int nrpy_x0t = Nxx0 >= %d ? min(nrpy_xt, Nxx0 - %d) : 0;  /* the tile, moved to fit */"
                    % (TX, TX),
                    ind + "const int i0 = NGHOSTS + nrpy_x0t + (int)threadIdx.x;",
                    ind + "const bool nrpy_own = nrpy_x0t + (int)threadIdx.x >= nrpy_xt;  /* else: the previous tile's point */",
                    ind + "const int nrpy_xl = (int)threadIdx.x;  /* lane within the row */",
                    ind + "const int nrpy_xm = min(%d, Nxx0);  /* lanes in the row */" % TX]
            continue
        if "const int x0_ = nrpy_bx * %d," % TX in line:      # the staged box starts at the moved tile
            out.append(line.replace("const int x0_ = nrpy_bx * %d," % TX, "const int x0_ = nrpy_x0t,"))
            continue
        if line.lstrip().startswith("if (%s && " % _SENR["even"]):  # 16-byte copy: also an even box start
            out.append(line.replace("if (%s && " % _SENR["even"], "if (%s && (x0_ & 1) == 0 && " % _SENR["even"]))
            continue
        sm = _STORE.match(line)
        if sm:
            out.append("%s*NRPY_OWNED(&%s[%s]) = %s;" % sm.group(1, 2, 3, 4))
            continue
        if line.strip().startswith("NRPY_RK4_STORE("):
            out.append(line.replace("NRPY_RK4_STORE(", "NRPY_RK4_STORE_OWNED("))
            continue
        m = _READ.match(line)
        if m and m.group(3) in q_of:
            gf, arr = m.group(3), m.group(2)
            q, H = q_of[gf], hx[gf]
            ind = re.match(r"^(\s*)", line).group(1)
            if q not in done:
                done.add(q)
                out.append("%sconst double nrpy_c%d = %s[IDX4(%s, i0,i1,i2)];" % (ind, q, arr, gf))
                out.append("%sconst double nrpy_h%d = %s[IDX4(%s, nrpy_xl < %d ? i0 - %d : i0 + %d,i1,i2)];"
                           % (ind, q, arr, gf, H, H, H))
            d = tuple(int(x) if x else 0 for x in m.group(4, 5, 6))
            if d == (0, 0, 0):
                out.append("%snrpy_c%d;" % (m.group(1), q))
                continue
            if not d[1] and not d[2]:
                moved.append((re.match(r"\s*const double (\w+) =", line).group(1),
                              "%snrpy_xshift<%d, %d, %d>(nrpy_c%d, nrpy_h%d, nrpy_xl, nrpy_xm);"
                              % (m.group(1), TX, H, d[0], q, q)))
                continue
        out.append(line)
    # each moved value is built right before the first statement that uses it
    pending = dict(moved)
    final = []
    for line in out:
        if pending:
            for name in [n for n in pending if re.search(r"\b%s\b" % n, line)]:
                final.append(pending.pop(name))
        final.append(line)
    if pending or any("i0" in l and "nrpy_bx*(int)blockDim.x" in l for l in final):
        return None
    return "\n".join(final)


# ---------------------------------------------------------------------------
# statement order: values built by cross-lane moves right before their first use
# ---------------------------------------------------------------------------
# The finite-difference emitter writes every stencil read and derivative first and
# the right-hand sides after them. The compiler spreads global loads towards their
# uses by itself, but not cross-lane moves: those stay where the source puts them,
# and a value built at the top stays live to its use. This pass moves the values
# built by nrpy_xshift, and the stencil-level definitions computed from them
# (derivatives, upwinded derivatives), down to just before the first statement of
# their block that uses them (the last statements first, so a moved definition's
# operands follow it down). Only these pure definitions move, and results are
# unchanged. (Moving every definition this way raises register use: the compiler's
# own order is better for the rest.)

_DEF = re.compile(r"^\s*const double (\w+) = [^;]*;$")
_IDENT = re.compile(r"[A-Za-z_]\w*")


def sink_xshift_values(kernel_src):
    lines = kernel_src.split("\n")
    block, bid = [], 0
    for line in lines:
        if "{" in line or "}" in line:
            bid += 1
        block.append(bid)
    defs = {}
    for k, line in enumerate(lines):
        m = _DEF.match(line)
        if m:
            defs[k] = m.group(1)
    # the x-values and every stencil-level definition (before the right-hand-side step)
    # computed from them, transitively
    rhs_step = min([k for k, line in enumerate(lines) if "evaluate right-hand sides" in line or
                    "Step 3 of 3" in line or "Step 2 of 2" in line] or [len(lines)])
    xvals = {defs[k] for k in defs if "nrpy_xshift<" in lines[k]}
    movable = set()
    for k in sorted(defs):
        if k < rhs_step and (defs[k] in xvals or xvals & set(_IDENT.findall(lines[k].split("=", 1)[1]))):
            movable.add(k)
            xvals.add(defs[k])
    uses = {}
    for k, line in enumerate(lines):
        for name in set(_IDENT.findall(line)):
            uses.setdefault(name, []).append(k)
    key = {k: float(k) for k in range(len(lines))}
    for k in sorted(movable, reverse=True):
        later = [u for u in uses.get(defs[k], []) if u > k]
        if not later:
            continue
        u = min(later, key=lambda x: key[x])
        if block[u] == block[k]:
            key[k] = key[u] - 1e-9 * (1 + k)
    return "\n".join(lines[k] for k in sorted(range(len(lines)), key=lambda k: key[k]))
