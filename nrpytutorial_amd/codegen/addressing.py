"""Tile-relative addressing of a kernel's gridfunction accesses.

A generated kernel reads `arr[IDX4(gf, i0+a, i1+b, i2+c)]`; IDX4 is a 64-bit index, so
every load gets its own 64-bit address in vector registers (two registers and a 64-bit
add per load). In a kernel whose block covers a fixed tile and whose accesses stay within
NGHOSTS of the thread's point, the index splits into

  * a part that is the same for the whole block: the gridfunction's offset and the
    tile's corner (64-bit, in scalar registers);
  * a 32-bit byte offset per lane (nrpy_off), measured from that corner moved back by
    NGHOSTS points in y and z (and in x for the "xl" form), so that it is never negative;
  * for the "xc" form, the x offset as a constant.

The loads then take the (scalar base + 32-bit lane offset + constant) form of the global
load instruction: one lane offset per stencil row serves every gridfunction. Two forms:

  "xc"  the x offset is the load's constant offset. One lane offset per (y, z) row; the
        compiler may fuse the loads of x-neighbours into 16-byte loads.
  "xl"  the x offset is in the lane offset: one lane offset per stencil point, no loads
        fused.

Which form is faster depends on the kernel's register pressure, so it is chosen per
kernel (TILE_ADDRESSING), from measurements. Values are unchanged: only the address
arithmetic differs.

The rewrite is a text pass over the kernel: it inserts NRPY_TILE_ORIGIN after the
definition of i2 and rebinds IDX4 to IDX4_TILE_XC / IDX4_TILE_XL (nrpy_runtime_pre.h)
inside the kernel; for the toolkit's access form, NRPY_ETK_TILE_ORIGIN and
CCTK_GFINDEX3D_TILE_XC / _XL. It refuses a kernel with an access it cannot prove in
range. The lane offset must stay below 2^32 bytes, which bounds the grid's xy-plane for
a given tile; the bundle's grid check rejects larger grids (TILE_BOUND).
"""
import os
import re

# kernel name (or a prefix ending in "*") -> "xc" or "xl"; NRPY_TILE_ADDR
# ("name=form,...", or "off") overrides for measurements
TILE_ADDRESSING = {}
# the bundle's NGHOSTS, the reach of every access the pass accepts
NGHOSTS = None
# largest tile (x, y, z) among the rewritten kernels; (0, 0, 0) while none is
TILE_BOUND = [0, 0, 0]
# kernel name -> form, of the kernels rewritten so far
APPLIED = {}


def form_for(name):
    env = os.environ.get("NRPY_TILE_ADDR")
    table = TILE_ADDRESSING
    if env is not None:
        if env in ("", "off"):
            return None
        table = dict(kv.split("=") for kv in env.split(","))
    if name in table:
        return table[name]
    for key, form in table.items():
        if key.endswith("*") and name.startswith(key[:-1]):
            return form
    return None


_I2 = re.compile(r"^\s*const int i2 = [^;]*\(int\)threadIdx\.z;$")
_I01 = (re.compile(r"^\s*const int i0 = [^;]*\(int\)threadIdx\.x;$", re.M),
        re.compile(r"^\s*const int i1 = [^;]*\(int\)threadIdx\.y;$", re.M))
# access macro -> (pattern of its uses, what its first argument must be (None: any),
# origin macro)
ACCESS = {"IDX4": (re.compile(r"\bIDX4S?\("), None, "NRPY_TILE_ORIGIN"),
          "CCTK_GFINDEX3D": (re.compile(r"\bCCTK_GFINDEX3D\("), "cctkGH", "NRPY_ETK_TILE_ORIGIN")}
# VJP_G_AT(k, o0, o1, o2) of the runtime expands to IDX4(k, (i0) + (o0), ...) (guarded
# by an interior test): accepted with literal offsets; the RK4 preload's
# IDX4(gf, i0, i1, i2) is the centre point and in range
_VJP = re.compile(r"\bVJP_G_AT\(")


def _args(src, start):
    """The comma-separated arguments of the call whose '(' is at src[start - 1]."""
    depth, args, cur = 1, [], ""
    i = start
    while depth:
        ch = src[i]
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
            if depth == 0:
                break
        if ch == "," and depth == 1:
            args.append(cur.strip())
            cur = ""
        else:
            cur += ch
        i += 1
    args.append(cur.strip())
    return args


def _offset(arg, var, nghosts):
    """a if arg is `var`, `var+a` or `var-a` with |a| <= nghosts, else None."""
    m = re.fullmatch(r"%s(?:\s*([+-])\s*(\d+))?" % var, arg)
    if not m:
        return None
    a = int(m.group(2) or 0) * (-1 if m.group(1) == "-" else 1)
    return a if abs(a) <= nghosts else None


def _x_ok(arg, nghosts):
    if _offset(arg, "i0", nghosts) is not None:
        return True
    # the halo load of a lane-move kernel: i0 - H in the row's first H lanes, else i0 + H
    m = re.fullmatch(r"nrpy_xl < (\d+) \? i0 - (\d+) : i0 \+ (\d+)", arg)
    return bool(m) and m.group(1) == m.group(2) == m.group(3) and int(m.group(1)) <= nghosts


def check_accesses(body, nghosts, access="IDX4"):
    """Raise ValueError unless every access through the access macro (IDX4 and IDX4S, or
    CCTK_GFINDEX3D) in the kernel body is within nghosts points of (i0, i1, i2) along
    each axis."""
    use, first, _ = ACCESS[access]
    if access != "IDX4" and _VJP.search(body):
        raise ValueError("tile addressing: VJP_G_AT in a kernel of the toolkit's access form")
    n = 0
    for m in _VJP.finditer(body):
        a = _args(body, m.end())
        if len(a) != 4 or not re.fullmatch(r"\d+", a[0]) or \
                not all(re.fullmatch(r"-?\d+", o) and abs(int(o)) <= nghosts for o in a[1:]):
            raise ValueError("tile addressing: VJP_G_AT(%s) not provably in range" % ", ".join(a))
        n += 1
    for m in use.finditer(body):
        a = _args(body, m.end())
        if len(a) != 4 or (first and a[0] != first) or not _x_ok(a[1], nghosts) \
                or _offset(a[2], "i1", nghosts) is None or _offset(a[3], "i2", nghosts) is None:
            raise ValueError("tile addressing: access %s(%s) not provably in range" %
                             (m.group(0)[:-1], ", ".join(a)))
        n += 1
    return n


def tile_address(kernel_src, form, tile, nghosts=None, name=None, access="IDX4"):
    """The kernel with tile-relative addressing of its accesses through the access macro
    (IDX4, or CCTK_GFINDEX3D), in form "xc" or "xl". tile is the kernel's fixed block
    shape, nghosts the bundle's NGHOSTS (default: NGHOSTS); raises ValueError if the
    kernel does not have the shape this relies on."""
    nghosts = NGHOSTS if nghosts is None else nghosts
    if not nghosts:
        raise ValueError("tile addressing: NGHOSTS unknown")
    if form not in ("xc", "xl"):
        raise ValueError("tile addressing form %r" % (form,))
    lines = kernel_src.split("\n")
    at = [k for k, l in enumerate(lines) if _I2.match(l)]
    if len(at) != 1 or not all(len(r.findall(kernel_src)) == 1 for r in _I01):
        raise ValueError("tile addressing: the kernel must define i0, i1, i2 once each, "
                         "ending in + threadIdx")
    close = max(k for k, l in enumerate(lines) if l == "}")
    body = "\n".join(lines[at[0] + 1:close])
    check_accesses(body, nghosts, access)
    origin = ACCESS[access][2]
    head = ["  %s(%s);" % (origin, "0" if form == "xc" else "NGHOSTS"),
            '#pragma push_macro("%s")' % access, "#undef %s" % access,
            "#define %s %s_TILE_%s" % (access, access, form.upper())]
    out = (lines[:at[0] + 1] + head + lines[at[0] + 1:close] + ['#pragma pop_macro("%s")' % access] +
           lines[close:])
    for d in range(3):
        TILE_BOUND[d] = max(TILE_BOUND[d], tile[d])
    if name:
        APPLIED[name] = form
    return "\n".join(out)


def bound_check_code(n0="p->Nxx_plus_2NGHOSTS0", n1="p->Nxx_plus_2NGHOSTS1"):
    """C condition under which a rewritten kernel's lane offset could reach 2^32 bytes, for
    rows of n0 and planes of n0 x n1 allocated points (C expressions), or None when no
    kernel is rewritten."""
    bx, by, bz = TILE_BOUND
    if not bx:
        return None
    return ("(size_t)(%d + 2*NGHOSTS) + (size_t)%s * ((%d + 2*NGHOSTS) + "
            "(size_t)%s * (%d + 2*NGHOSTS)) > ((size_t)1 << 29)" % (bx, n0, by, n1, bz))
