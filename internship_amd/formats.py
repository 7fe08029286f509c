"""The reference's interchange formats, read and written in this package's
own code.

- `NNWeight.cuh` (core/NNWeight.cuh): the MRPNN weights as C arrays,
  `__device__ __constant__ float NAME[N] = {v,v,...};` (the 16 large arrays
  as `__device__ const float`), values in shortest round-trip FP32 text.
  write_nnweight_header exports any weight vector of that layout (for example
  the natively trained network) so the reference can compile it in;
  read_nnweight_header parses one (the parser of gen_tables.py).
- `MinLossWeights.txt` (train/train.py:320-355): `const float NAME[N] = {`,
  then 10 values per line with 8 decimals, `};` and a blank line per array.
- The header `tools/CastWeight.py` makes of a MinLossWeights.txt
  (tools/CastWeight.py:22-28): `__constant__ half NAME[N];` declarations,
  host initialisers `half NAME_[N] = {...}` of the same decimal text with an
  `f` suffix, and a `SetWeights()` that copies them to the symbols.
- The training-data CSV of dataGen.cu (dataGen.cu:316-332): one `#` comment
  header line, then one sample per row, fixed-point values with 6 decimals
  separated by commas.  The reference's rows hold the RPNN descriptor
  (10 x 225 densities, 10 x 225 variances, gamma, radiance); this package's
  rows hold the MRPNN sample (576 features, gamma, g, srp, radiance) with the
  same conventions.
- Volume files (core/volume.cu:835-911): `.bin` = int32 resolution + res^3
  FP32 (z fastest); `.txt` = the resolution then res^3 values as text;
  `.vox` = 256^3 byte values as text, 64 per line after `w`/`h`/`d` header
  lines, scaled by 64/255.  The C++ facade reads the same files.
"""
import re

import numpy as np

from . import weight_manifest

# ----------------------------------------------------------------- weights


def _arrays(weights):
    w = np.asarray(weights, np.float32).ravel()
    man = weight_manifest()
    if w.size != man["total"]:
        raise ValueError("expected %d weights, got %d" % (man["total"], w.size))
    return [(a["name"], w[a["offset"]:a["offset"] + a["count"]], a.get("storage", "constant")) for a in man["arrays"]]


def _from_named(named):
    """{name: values} -> the 49,673-value vector in manifest order (every array, exact counts)."""
    man = weight_manifest()
    out = np.zeros(man["total"], np.float32)
    for a in man["arrays"]:
        v = named.get(a["name"])
        if v is None:
            raise ValueError("missing weight array %s" % a["name"])
        if len(v) != a["count"]:
            raise ValueError("%s: %d values, expected %d" % (a["name"], len(v), a["count"]))
        out[a["offset"]:a["offset"] + a["count"]] = np.asarray(v, np.float32)
    return out


def write_nnweight_header(weights, path):
    lines = ["#pragma once"]
    for name, vals, storage in _arrays(weights):
        qual = "__device__ __constant__ float" if storage == "constant" else "__device__ const float"
        lines.append("%s %s[%d] = {" % (qual, name, len(vals)))
        lines.append(",".join(str(np.float32(v)) for v in vals))
        lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_nnweight_header(path):
    from .gen_tables import parse_weights
    return _from_named({n: v for n, v in parse_weights(open(path).read())})


def write_min_loss_weights(weights, path):
    with open(path, "w") as f:
        for name, vals, _ in _arrays(weights):
            f.write("const float %s[%d] = {\n" % (name, len(vals)))
            for i in range(0, len(vals), 10):
                f.write("    " + ", ".join("%.8f" % float(v) for v in vals[i:i + 10]))
                f.write(",\n" if i + 10 < len(vals) else "\n")
            f.write("};\n\n")


_MLW = re.compile(r"const float (\w+)\[(\d+)\] = \{\n(.*?)\n\};", re.S)


def read_min_loss_weights(path):
    named = {}
    for name, n, body in _MLW.findall(open(path).read()):
        vals = [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]
        if len(vals) != int(n):
            raise ValueError("%s: %d values, declared %s" % (name, len(vals), n))
        named[name] = vals
    return _from_named(named)


def write_castweight_header(min_loss_weights_text):
    """The header tools/CastWeight.py generates from a MinLossWeights.txt text
    (FP16 constants filled by SetWeights())."""
    device, host, code = "", "", "void SetWeights()\n{\n"
    for name, n, body in _MLW.findall(min_loss_weights_text):
        value = (body + ",").replace(" ", "").replace(",", "f,")
        device += "__constant__ half %s[%d];\n" % (name, int(n))
        host += "half %s_[%d] = {\n%s\n};\n" % (name, int(n), value)
        code += "    cudaMemcpyToSymbol(%s, %s_, %d*sizeof(half));\n" % (name, name, int(n))
    code += "}"
    sep = "\n//-----------------------------------------\n\n"
    return ("#pragma once\n#include <cuda_runtime.h>\n#include <cuda_fp16.h>\n\n" + sep + device + "\n" + sep[1:] + host
            + "\n" + sep[1:] + code)


_CAST = re.compile(r"half (\w+)_\[(\d+)\] = \{\n(.*?)\n\};", re.S)


def read_castweight_header(text):
    """Weights of a CastWeight.py header as its `half` initialisers store them:
    each `...f` literal rounded to FP32, then to FP16."""
    named = {}
    for name, n, body in _CAST.findall(text):
        vals = [float(v.rstrip("f")) for v in body.replace("\n", "").split(",") if v.strip()]
        if len(vals) != int(n):
            raise ValueError("%s: %d values, declared %s" % (name, len(vals), n))
        # `half x = v f`: the literal is an FP32 value, then rounded to FP16
        named[name] = np.asarray(vals, np.float64).astype(np.float32).astype(np.float16).astype(np.float32)
    return _from_named(named)


# -------------------------------------------------------- training samples

def write_training_csv(path, feat, aux):
    """MRPNN samples (feat (n,576), aux (n,4) = gamma, g, srp, radiance) in
    dataGen's CSV conventions."""
    feat = np.asarray(feat, np.float32)
    aux = np.asarray(aux, np.float32)
    with open(path, "w") as f:
        f.write("# %dx(576 MRPNN features + gamma, g, srp, radiance)\n" % len(feat))
        rows = np.concatenate([feat, aux], axis=1)
        np.savetxt(f, rows, fmt="%.6f", delimiter=",")


def read_training_csv(path):
    """(header, feat (n,576), aux (n,4)) of a CSV written by write_training_csv."""
    with open(path) as f:
        header = f.readline().rstrip("\n")
        rows = np.loadtxt(f, delimiter=",", dtype=np.float64, ndmin=2)
    if not header.startswith("#"):
        raise ValueError("no '#' header line")
    if rows.size and rows.shape[1] != 580:
        raise ValueError("rows have %d columns, expected 580" % rows.shape[1])
    rows = rows.reshape(-1, 580).astype(np.float32)
    return header, rows[:, :576], rows[:, 576:]


# ----------------------------------------------------------------- volumes

def write_volume_bin(path, grid):
    """`path` is the file name itself (the reference appends `.bin` to its argument)."""
    g = np.ascontiguousarray(grid, np.float32)
    if g.ndim != 3 or len(set(g.shape)) != 1:
        raise ValueError("grid must be res^3")
    with open(path, "wb") as f:
        f.write(np.int32(g.shape[0]).tobytes())
        f.write(g.tobytes())


def write_volume_txt(path, grid):
    g = np.asarray(grid, np.float32)
    with open(path, "w") as f:
        f.write("%d\n" % g.shape[0])
        np.savetxt(f, g.reshape(-1, g.shape[0]), fmt="%.9g")


def read_volume(path):
    """The grid (res^3 FP32, z fastest) of `path + ".bin"` if that exists, else
    of `path` as `.txt` or `.vox` (core/volume.cu:835-911)."""
    import os
    if os.path.exists(path + ".bin"):
        raw = open(path + ".bin", "rb").read()
        res = int(np.frombuffer(raw[:4], "<i4")[0])
        if not 0 < res <= 1024 or len(raw) < 4 + 4 * res ** 3:
            raise ValueError("bad .bin volume")
        return np.frombuffer(raw[4:4 + 4 * res ** 3], "<f4").reshape(res, res, res).copy()
    if path.endswith("txt"):
        vals = open(path).read().split()
        res = int(vals[0])
        g = np.zeros(res ** 3, np.float32)
        data = np.asarray(vals[1:1 + res ** 3], np.float32)
        g[:len(data)] = data
        return g.reshape(res, res, res)
    if path.endswith("vox"):
        g = np.zeros(256 ** 3, np.float32)
        i = 0
        for line in open(path):
            if not line.strip() or line[0] in "whd":
                continue
            vals = [int(v) for v in line.split()[:64]]
            n = min(len(vals), g.size - i)
            g[i:i + n] = np.asarray(vals[:n], np.float32) * np.float32(64.0 / 255.0)
            i += n
            if i >= g.size:
                break
        return g.reshape(256, 256, 256)
    raise ValueError("unknown volume format: %s" % path)
