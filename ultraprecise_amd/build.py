"""Build step for the HIP library: embeds the device headers for hipRTC and
compiles libupdec.so for gfx950 in-tree (the built library travels with the
source tree; nothing is installed)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libupdec.so")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


SOURCES = ("updec_api.hip", "vm.hip", "codec.hip", "segsum.hip", "planner.cpp")


def embed_device_header():
    """dec_device.h and dec_mlane.h as string constants, the headers hipRTC
    is given for every generated kernel"""
    delim = "UPDEC_EMBED"
    out = "// generated from dec_device.h and dec_mlane.h by build.py; do not edit\n"
    for header, name in (("dec_device.h", "kDecDeviceSource"), ("dec_mlane.h", "kDecMlaneSource")):
        src = open(os.path.join(CSRC, header)).read()
        assert ")" + delim + '"' not in src
        out += "static const char " + name + "[] = R\"" + delim + "(" + src + ")" + delim + "\";\n"
    path = os.path.join(CSRC, "dec_device_embed.inc")
    if not os.path.exists(path) or open(path).read() != out:
        with open(path, "w") as f:
            f.write(out)


def build(verbose=False):
    embed_device_header()
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.rsplit(".", 1)[0] + ".o")
        cmd = [hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [hipcc, "-x", "hip"] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(obj)
    # the translation units are independent: compile them side by side
    procs = [subprocess.Popen(c) for c in cmds]
    codes = [p.wait() for p in procs]
    for c, rc in zip(cmds, codes):
        if rc:
            raise subprocess.CalledProcessError(rc, c)
    tmp = LIB + ".tmp"
    cmd = [hipcc, "-shared", "-fPIC", "--offload-arch=gfx950"] + objs + ["-lhiprtc", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
