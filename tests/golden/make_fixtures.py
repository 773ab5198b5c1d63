"""Convert the reference's golden convolution vectors into repo-local fixtures.

Source: the reference's test resources
  src/test/resources/platform/opencl/hyperkernels/convolutionDynamic/
read by src/test/scala/cogx/DynamicConvolutionSpec.scala:131-197 (x, k, wx, y,
wty and the .hdr ConvolveOp headers, L40-75) with the field reader of
src/test/scala/cogx/helper/fieldio/FieldIO.scala:75-140.

.fld format (java.io.DataOutputStream, big-endian): int n, n UTF-16 chars of the
tensor kind, int dims, dims ints (rows, cols), int tensorDims, tensorDims ints,
long tick, then float32 values row-major.  .hdr: int n, n chars of the sampling
policy name, int step, int m, m chars of the border policy name.

Output (little-endian, numpy .npz without pickles):
  tests/golden/dyn_<i>.npz        x, k, wx, y, wty of golden case i (0..24)
  tests/golden/orphan_<t>.npz     target t, filter t, and the four edge bands
                                  (20 rows/cols) plus a 32x32 centre block of
                                  convolve{Zero,Clamp,Cyclic}<t>
  tests/golden/manifest.json      border, sampling, step and shapes per case

Run in the build container (the reference tree is not present on GPU hosts):
  python tests/golden/make_fixtures.py /root/reference
"""
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REL = "src/test/resources/platform/opencl/hyperkernels/convolutionDynamic"
ORPHANS = (0, 3)
BAND = 20


def read_fld(path):
    b = open(path, "rb").read()
    o = 0

    def take(fmt):
        nonlocal o
        v = struct.unpack_from(fmt, b, o)
        o += struct.calcsize(fmt)
        return v

    (n,) = take(">i")
    kind = b[o:o + 2 * n].decode("utf-16-be")
    o += 2 * n
    (d,) = take(">i")
    dims = take(">%di" % d)
    (t,) = take(">i")
    tdims = take(">%di" % t) if t else ()
    take(">q")  # tick
    if kind != "ScalarLattice" or tdims:
        raise ValueError("%s: only scalar fields are used (%s %s)" % (path, kind, tdims))
    count = int(np.prod(dims))
    data = np.frombuffer(b, dtype=">f4", count=count, offset=o).astype("<f4")
    return data.reshape(dims)


def read_hdr(path):
    b = open(path, "rb").read()
    (n,) = struct.unpack_from(">i", b, 0)
    sampling = b[4:4 + 2 * n].decode("utf-16-be")
    o = 4 + 2 * n
    (step,) = struct.unpack_from(">i", b, o)
    o += 4
    (m,) = struct.unpack_from(">i", b, o)
    o += 4
    border = b[o:o + 2 * m].decode("utf-16-be")
    return sampling, step, border


def main(ref_root):
    src = os.path.join(ref_root, REL)
    manifest = {"source": REL, "dynamic": [], "orphan": []}
    for i in range(25):
        sampling, step, border = read_hdr(os.path.join(src, "%d.hdr" % i))
        arrs = {n: read_fld(os.path.join(src, "%s%d.fld" % (n, i))) for n in ("x", "k", "wx", "y", "wty")}
        np.savez(os.path.join(HERE, "dyn_%d.npz" % i), **arrs)
        manifest["dynamic"].append({
            "case": i, "sampling": sampling, "step": step, "border": border,
            "shapes": {n: list(a.shape) for n, a in arrs.items()},
            "files": ["%s%d.fld" % (n, i) for n in ("x", "k", "wx", "y", "wty")] + ["%d.hdr" % i]})
    for t in ORPHANS:
        target = read_fld(os.path.join(src, "target%d.fld" % t))
        filt = read_fld(os.path.join(src, "filter%d.fld" % t))
        arrs = {"target": target, "filter": filt}
        H, W = target.shape
        r0, c0 = H // 2 - 16, W // 2 - 16
        for border in ("Zero", "Clamp", "Cyclic"):
            out = read_fld(os.path.join(src, "convolve%s%d.fld" % (border, t)))
            assert out.shape == target.shape
            arrs[border + "_top"] = out[:BAND]
            arrs[border + "_bottom"] = out[-BAND:]
            arrs[border + "_left"] = out[:, :BAND]
            arrs[border + "_right"] = out[:, -BAND:]
            arrs[border + "_centre"] = out[r0:r0 + 32, c0:c0 + 32]
        np.savez(os.path.join(HERE, "orphan_%d.npz" % t), **arrs)
        manifest["orphan"].append({"case": t, "target": list(target.shape), "filter": list(filt.shape),
                                   "band": BAND, "centre_origin": [r0, c0],
                                   "files": ["target%d.fld" % t, "filter%d.fld" % t] +
                                            ["convolve%s%d.fld" % (b, t) for b in ("Zero", "Clamp", "Cyclic")]})
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
