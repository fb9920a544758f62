"""Build every system bundle for gfx950 (and, for tests, the C oracle).

1. Generate ``csrc/generated/<bundle>.hip`` through the codegen API, one fresh
   interpreter per bundle; skipped when the stamp (hash of the generator
   sources) is unchanged, because BSSN codegen takes minutes of SymPy CSE.
2. Compile each bundle with hipcc into ``lib/libnrpy_<bundle>.so``; this
   always runs.

Usage: python -m nrpytutorial_amd.build [bundle ...]
"""
import concurrent.futures as cf
import glob
import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
GENDIR = os.path.join(CSRC, "generated")
LIBDIR = os.path.join(HERE, "lib")
ORACLE = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-variable", "-Wno-unused-but-set-variable"]

from .codegen.systems import BUNDLES  # noqa: E402


def _sources_hash(extra=()):
    h = hashlib.sha1()
    files = sorted(glob.glob(os.path.join(HERE, "**", "*.py"), recursive=True))
    files = [f for f in files if os.path.basename(f) != "build.py" and "/lib/" not in f]
    for f in list(files) + list(extra):
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def _stale(stamp_path, digest, outputs):
    if not all(os.path.exists(o) for o in outputs):
        return True
    if not os.path.exists(stamp_path):
        return True
    with open(stamp_path) as f:
        return f.read().strip() != digest


def generate(names, force=False, jobs=4):
    digest = _sources_hash()
    todo = []
    for n in names:
        outs = [os.path.join(GENDIR, n + ".hip"), os.path.join(GENDIR, n + ".json")]
        if force or _stale(os.path.join(GENDIR, n + ".stamp"), digest, outs):
            todo.append(n)

    def run(n):
        r = subprocess.run([sys.executable, "-m", "nrpytutorial_amd.codegen.systems", n], cwd=ROOT,
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("codegen of %s failed:\n%s\n%s" % (n, r.stdout, r.stderr))
        with open(os.path.join(GENDIR, n + ".stamp"), "w") as f:
            f.write(digest)
        return n
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for n in ex.map(run, todo):
            print("generated", n, flush=True)


def compile_bundle(n):
    os.makedirs(LIBDIR, exist_ok=True)
    src = os.path.join(GENDIR, n + ".hip")
    out = os.path.join(LIBDIR, "libnrpy_%s.so" % n)
    tmp = out + ".tmp"
    r = subprocess.run([HIPCC] + HIP_FLAGS + [src, "-o", tmp], cwd=CSRC, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (n, r.stderr[-20000:]))
    os.replace(tmp, out)
    return out


def build_oracle(names, force=False, jobs=4):
    """Generate and gcc-compile the oracle (test infrastructure)."""
    gen = os.path.join(ORACLE, "generated")
    digest = _sources_hash(extra=[os.path.join(ORACLE, "gen_oracle.py")])
    todo = [n for n in names
            if force or _stale(os.path.join(gen, n + ".stamp"), digest, [os.path.join(gen, n + ".c")])]

    def run(n):
        r = subprocess.run([sys.executable, os.path.join(ORACLE, "gen_oracle.py"), n], cwd=ROOT,
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("oracle codegen of %s failed:\n%s\n%s" % (n, r.stdout, r.stderr))
        with open(os.path.join(gen, n + ".stamp"), "w") as f:
            f.write(digest)
        return n
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(run, todo))
    r = subprocess.run(["make", "-C", ORACLE, "-j8", "BUNDLES=" + " ".join(names)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout[-5000:] + r.stderr[-5000:])


def build_all(names=None, oracle=True, force=False):
    names = list(names or BUNDLES)
    generate(names, force=force)
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        for out in ex.map(compile_bundle, names):
            print("compiled", out, flush=True)
    if oracle:
        build_oracle([n for n in ORACLE_BUNDLES if ORACLE_BUNDLES[n] in names], force=force)


# oracle library -> the product bundle it checks ("_ld": the extended-precision build)
ORACLE_BUNDLES = {n: n for n in ("scalarwave_cartesian", "bssn_cartesian", "bssn_spherical_sym2",
                                 "scalarwave_sinhsphericalv2", "scalarwave_sinhcylindricalv2", "maxwell_cartesian",
                                 "scalarwave_symtp", "scalarwave_sinhspherical", "scalarwave_sinhcylindrical",
                                 "scalarwave_sinhsymtp", "maxwell_sinhcylindrical")}
ORACLE_BUNDLES["bssn_cartesian_ld"] = "bssn_cartesian"

if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("-")]
    build_all(args or None, oracle="--no-oracle" not in sys.argv, force="--force" in sys.argv)
