"""Build of ``libigaops_hip.so`` (all HIP sources of ``csrc/`` for gfx950), in-tree.

``python -m igaops_amd.build`` or ``__graft_entry__.build()``.  The library lands
next to this file so that it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libigaops_hip.so"
INCLUDE = PKG.parent / "include"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(CSRC.glob("*.hip"))


def build(verbose: bool = True, jobs: int = 8) -> Path:
    """Compile every ``csrc/*.hip`` for gfx950 and link the shared library."""
    obj_dir = PKG / "_build"
    obj_dir.mkdir(exist_ok=True)
    flags = [
        "-O3",
        "-std=c++17",
        "--offload-arch=gfx950",
        "-fPIC",
        "-Wall",
        "-Wno-unused-result",
        # occupancy hints that LDS use overrides for the widest bands are expected
        "-Wno-pass-failed",
        # packed FP32 instructions are kept off in all device code
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        f"-I{INCLUDE}",
    ]
    objs = []
    procs = []
    for src in sources():
        obj = obj_dir / (src.stem + ".o")
        objs.append(obj)
        cmd = [hipcc(), *flags, "-c", str(src), "-o", str(obj)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((cmd, text))
        elif verbose and text.strip():
            # keep warnings visible, but drop the known host-pass notice about the feature flag
            lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l]
            if lines:
                print("\n".join(lines), file=sys.stderr)
    if failed:
        msg = "\n\n".join(" ".join(c) + "\n" + t for c, t in failed)
        raise RuntimeError("HIP compilation failed:\n" + msg)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + res.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
