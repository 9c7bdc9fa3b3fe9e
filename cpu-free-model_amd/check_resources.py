"""Register and scratch check of every gfx950 kernel in libcfm.

Compiles each HIP source of the library for the device only with the
compiler's kernel-resource-usage remarks and fails (exit 1) when any kernel
uses scratch memory (per-lane stack: VGPR spills or a dynamic stack).  SGPR
spills, which go to VGPR lanes and not to memory, are listed but allowed.

    python cpu-free-model_amd/check_resources.py [--list]

The flags are the library's own (Makefile HIPFLAGS); objects go to a
temporary directory.
"""
import argparse
import concurrent.futures as cf
import os
import re
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["api.hip", "jacobi.hip", "jacobi_noalloc.hip", "cg.hip", "cg_batch.hip", "unet.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Xclang", "-target-feature",
         "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(REPO, "include")]


def remarks(src, tmp):
    out = os.path.join(tmp, os.path.basename(src) + ".o")
    r = subprocess.run([HIPCC] + FLAGS + ["--offload-device-only", "-Rpass-analysis=kernel-resource-usage", "-c",
                                          os.path.join(PKG, "csrc", src), "-o", out],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{src}: compile failed\n{r.stderr[-2000:]}")
    return r.stderr


def parse(text):
    kernels, cur = [], None
    for ln in text.splitlines():
        m = re.search(r"remark: Function Name: (\S+)", ln)
        if m:
            cur = {"name": m.group(1)}
            kernels.append(cur)
            continue
        m = re.search(r"remark:\s+([A-Za-z ]+?)(?: \[[^\]]*\])?: (\S+) \[-Rpass", ln)
        if m and cur is not None:
            cur[m.group(1).strip()] = m.group(2)
    return kernels


def demangle(names):
    try:
        r = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True, check=True)
        return r.stdout.splitlines()
    except (OSError, subprocess.CalledProcessError):
        return names


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--list", action="store_true", help="print every kernel's VGPRs, spills and occupancy")
    args = ap.parse_args()
    with tempfile.TemporaryDirectory() as tmp, cf.ThreadPoolExecutor(len(SOURCES)) as ex:
        texts = dict(zip(SOURCES, ex.map(lambda s: remarks(s, tmp), SOURCES)))
    bad, total = [], 0
    for src, text in texts.items():
        ks = parse(text)
        total += len(ks)
        for k, dn in zip(ks, demangle([k["name"] for k in ks])):
            scratch = int(k.get("ScratchSize", "0"))
            dyn = k.get("Dynamic Stack", "False") == "True"
            if args.list:
                print(f"{src:13s} vgpr {k.get('VGPRs', '?'):>3s} agpr {k.get('AGPRs', '?'):>3s} "
                      f"sgpr-spill {k.get('SGPRs Spill', '?'):>3s} vgpr-spill {k.get('VGPRs Spill', '?'):>3s} "
                      f"scratch {scratch:4d} occ {k.get('Occupancy', '?')}  {dn}")
            if scratch > 0 or dyn:
                bad.append(f"{src}: {dn}: scratch {scratch} B/lane, VGPR spills {k.get('VGPRs Spill')}, "
                           f"dynamic stack {dyn}")
    if total == 0:
        print("check_resources: no kernel remarks found", file=sys.stderr)
        return 1
    if bad:
        print("check_resources: kernels using scratch memory:\n  " + "\n  ".join(bad), file=sys.stderr)
        return 1
    print(f"check_resources: {total} kernels, none uses scratch memory")
    return 0


if __name__ == "__main__":
    sys.exit(main())
