"""Build the native library ``libpetgem_amd.so`` in-tree with hipcc for gfx950.

Usage: ``python -m petgem_amd.build`` (or ``__graft_entry__.build()``).
Every HIP/C++ source under ``petgem_amd/csrc`` is compiled to an object and
linked into one shared library next to this file.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpetgem_amd.so")
SOURCES = ["pg_api.cpp", "pg_plan.cpp", "pg_element.hip", "pg_krylov.hip", "pg_batch.hip",
           "pg_multi.hip", "pg_elemop.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")]


# per-source flags: the element products at p = 3 unroll 1035 entry pairs, past
# the default size limit of a fully unrolled loop
EXTRA = {"pg_elemop.hip": ["-mllvm", "-pragma-unroll-threshold=2000000"]}


def _stale(obj, src):
    """True when obj is missing or older than its source or any header it may include."""
    if not os.path.exists(obj):
        return True
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC)
                                        if h.endswith(".h")]
    deps.append(os.path.join(HERE, "..", "include", "petgem_amd.h"))
    return any(os.path.getmtime(d) > os.path.getmtime(obj) for d in deps)


def build(verbose=False, jobs=4, incremental=False):
    """Compile every source (incremental=True: only those whose object is stale) and link."""
    os.makedirs(BUILD, exist_ok=True)
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        if incremental and not _stale(obj, src):
            objs.append(obj)
            continue
        cmd = [HIPCC] + FLAGS + EXTRA.get(src, []) + ["-x", "hip", "-c", os.path.join(CSRC, src),
                                                      "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, pr in procs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            failed.append((src, out.decode(errors="replace")))
        elif verbose and out:
            print(out.decode(errors="replace"))
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
