"""Register and scratch report of the built gfx950 kernels.

The Makefile passes -Rpass-analysis=kernel-resource-usage to hipcc, whose
remarks (one block per kernel: VGPRs, AGPRs, ScratchSize, Occupancy) go to
_build/resources.log.  check() summarises them and fails when an Euler or
diffusion kernel uses scratch (register spills or a private array indexed at
run time): the flux kernels are laid out so that their working set stays in
registers, and a kernel that spills has regressed.
"""
import re
import sys

NO_SCRATCH = ("euler_", "face_divergence", "diffusion_")
FIELDS = ("VGPRs", "AGPRs", "ScratchSize [bytes/lane]", "Occupancy [waves/SIMD]")


def parse(text):
    out = {}
    for block in re.split(r"remark: [^\n]*?Function Name: ", text)[1:]:
        name = block.split()[0]
        vals = {}
        for k in FIELDS:
            m = re.search(re.escape(k) + r": (\d+)", block)
            vals[k] = int(m.group(1)) if m else -1
        out[name] = vals
    return out


def check(path, verbose=False):
    with open(path) as fh:
        kernels = parse(fh.read())
    bad = []
    for name, v in sorted(kernels.items()):
        if verbose:
            print("%-90s vgpr %3d agpr %3d scratch %4d occ %d" % (name[:90], v["VGPRs"], v["AGPRs"],
                  v["ScratchSize [bytes/lane]"], v["Occupancy [waves/SIMD]"]))
        if any(k in name for k in NO_SCRATCH) and v["ScratchSize [bytes/lane]"] > 0:
            bad.append((name, v["ScratchSize [bytes/lane]"]))
    if bad:
        raise RuntimeError("flux kernels using scratch: " + ", ".join("%s (%d B/lane)" % b for b in bad))
    return kernels


if __name__ == "__main__":
    check(sys.argv[1], verbose=True)
