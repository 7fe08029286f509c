"""The reference's interchange formats (internship_amd/formats.py and the C++
facade's volume loader), on the CPU:

- NNWeight.cuh (core/NNWeight.cuh): the shipped and the natively trained
  weights written and parsed back bit for bit, with the reference's storage
  classes (130 `__constant__`, 16 global `const` arrays);
- MinLossWeights.txt (train/train.py:320-355): 10 values per line with 8
  decimals; parsed back within half a unit in the 8th decimal;
- the CastWeight.py header (tools/CastWeight.py:22-28): one FP16 constant and
  one host initialiser per array, SetWeights() with one copy per array; its
  values are the FP16 rounding of the MinLossWeights text;
- dataGen's CSV (dataGen.cu:316-332): a '#' header, 6-decimal rows;
- volume files (core/volume.cu:835-911): .bin / .txt / .vox written here and
  read back by formats.read_volume and by the facade's LoadVolumeFile, which
  must agree value for value.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import internship_amd as ia
from internship_amd import formats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("which", ["shipped", "native"])
def test_nnweight_header_round_trip(tmp_path, which):
    w = ia.load_reference_weights() if which == "shipped" else ia.load_native_weights()
    p = tmp_path / "NNWeight.cuh"
    formats.write_nnweight_header(w, str(p))
    text = p.read_text()
    assert text.count("__device__ __constant__ float") == 130 and text.count("__device__ const float") == 16
    assert "__device__ const float LC0W[16384] = {" in text
    assert np.array_equal(formats.read_nnweight_header(str(p)), w)


def test_min_loss_weights_round_trip(tmp_path):
    w = ia.load_native_weights()
    p = tmp_path / "MinLossWeights.txt"
    formats.write_min_loss_weights(w, str(p))
    lines = p.read_text().split("\n")
    assert lines[0] == "const float LD01W[64] = {"
    assert len(lines[1].split(",")) == 11 and lines[1].startswith("    ")  # 10 values and a trailing comma
    assert all(len(v.strip().split(".")[1]) == 8 for v in lines[1].rstrip(",").split(","))
    back = formats.read_min_loss_weights(str(p))
    assert np.abs(back.astype(np.float64) - w).max() <= 5.0000001e-9 + 1e-7 * np.abs(w).max()


def test_castweight_header(tmp_path):
    w = ia.load_reference_weights()
    p = tmp_path / "MinLossWeights.txt"
    formats.write_min_loss_weights(w, str(p))
    h = formats.write_castweight_header(p.read_text())
    assert h.startswith("#pragma once\n#include <cuda_runtime.h>\n#include <cuda_fp16.h>")
    assert h.count("__constant__ half ") == 146 and h.count("cudaMemcpyToSymbol(") == 146
    assert "__constant__ half LX12B[1];" in h and "cudaMemcpyToSymbol(LX12B, LX12B_, 1*sizeof(half));" in h
    back = formats.read_castweight_header(h)
    mlw = formats.read_min_loss_weights(str(p))
    assert np.array_equal(back, mlw.astype(np.float16).astype(np.float32))


def test_training_csv_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    feat = rng.uniform(0, 3, (40, 576)).astype(np.float32)
    aux = np.stack([rng.uniform(0, 3, 40), np.full(40, 0.8), np.full(40, 0.96), rng.uniform(0, 2, 40)],
                   axis=1).astype(np.float32)
    p = tmp_path / "DS.csv"
    formats.write_training_csv(str(p), feat, aux)
    header, f2, a2 = formats.read_training_csv(str(p))
    assert header.startswith("# 40x")
    row = p.read_text().split("\n")[1].split(",")
    assert len(row) == 580 and all(len(v.split(".")[1]) == 6 for v in row)
    assert np.abs(f2 - feat).max() <= 5.1e-7 * 4 and np.abs(a2 - aux).max() <= 5.1e-7 * 4


def _facade_loader(tmp_path):
    src = tmp_path / "load.cpp"
    src.write_text(r'''
#include <cstdio>
#include "mrp_volume_render.hpp"
int main(int argc, char** argv) {
    int res = 0; bool bin = false; std::vector<float> d;
    if (!internship::LoadVolumeFile(argv[1], &res, &d, &bin)) return 1;
    FILE* f = std::fopen(argv[2], "wb");
    std::fwrite(&res, 4, 1, f);
    std::fwrite(d.data(), 4, d.size(), f);
    std::fclose(f);
    return bin ? 10 : 0;
}
''')
    exe = tmp_path / "load"
    gxx = shutil.which("g++") or "g++"
    r = subprocess.run([gxx, "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

    def load(path):
        out = tmp_path / "out.raw"
        rc = subprocess.run([str(exe), path, str(out)]).returncode
        raw = out.read_bytes()
        res = int(np.frombuffer(raw[:4], "<i4")[0])
        return rc, np.frombuffer(raw[4:], "<f4").reshape(res, res, res)
    return load


def test_volume_files_bin_txt_vox(tmp_path):
    load = _facade_loader(tmp_path)
    rng = np.random.default_rng(4)
    g = rng.random((12, 12, 12)).astype(np.float32)
    # .bin: the constructor's argument is the path without the extension
    formats.write_volume_bin(str(tmp_path / "a.bin"), g)
    assert np.array_equal(formats.read_volume(str(tmp_path / "a")), g)
    rc, fg = load(str(tmp_path / "a"))
    assert rc == 10 and np.array_equal(fg, g)
    # .txt
    formats.write_volume_txt(str(tmp_path / "b.txt"), g)
    assert np.array_equal(formats.read_volume(str(tmp_path / "b.txt")), g)
    rc, fg = load(str(tmp_path / "b.txt"))
    assert rc == 0 and np.array_equal(fg, g)
    # .vox: 256^3 bytes as text, 64 per line, after w/h/d header lines
    v = rng.integers(0, 256, 256 ** 3)
    with open(tmp_path / "c.vox", "w") as f:
        f.write("w 256\nh 256\nd 256\n")
        np.savetxt(f, v.reshape(-1, 64), fmt="%d")
    want = v.astype(np.float32) * np.float32(64.0 / 255.0)
    got = formats.read_volume(str(tmp_path / "c.vox"))
    assert np.array_equal(got.ravel(), want)
    rc, fg = load(str(tmp_path / "c.vox"))
    assert rc == 0 and np.array_equal(fg.ravel(), want)
