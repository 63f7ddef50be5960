"""The drop-in apps end to end on the GPU: METIS file in, one block id per line
out (the reference's app/GPU-HeiPa.cpp and app/GPU-HeiProMap.cpp surface),
bit-exact with the oracle for a fixed seed."""
import os
import subprocess

import numpy as np
import pytest

import heipa
import heipa_oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "gpu-heipa_amd", "bin")


def _run(args):
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_gpu_heipa_cli(tmp_path):
    g = heipa.gen_rgg(5000, 10.0, 4, 10)
    path = str(tmp_path / "rgg.graph")
    heipa.write_metis(path, g)
    out = str(tmp_path / "part.txt")
    _run([os.path.join(BIN, "GPU-HeiPa"), "--graph", path, "--k", "8", "--imbalance", "0.03", "--config",
          "ultra", "--seed", "1", "--mapping", out, "--verbose-level", "0", "--arena-gb", "1"])
    part = np.loadtxt(out, dtype=np.uint32)
    ref, cut, maxw, _ = orc.partition(g, 8, 0.03, 1, True)
    assert np.array_equal(part, ref)
    assert orc.edge_cut(g, part) == cut


def test_gpu_heipromap_cli(tmp_path):
    g = heipa.gen_rmat(12, 0, 8.0, 0.57, 0.19, 0.19, 6, 100, True)
    path = str(tmp_path / "rmat.graph")
    heipa.write_metis(path, g)
    out = str(tmp_path / "map.txt")
    _run([os.path.join(BIN, "GPU-HeiProMap"), "--graph", path, "--hierarchy", "2:4:3", "--distance", "1:10:100",
          "--imbalance", "0.03", "--config", "HM", "--seed", "1", "--mapping", out, "--verbose-level", "0",
          "--arena-gb", "1"])
    part = np.loadtxt(out, dtype=np.uint32)
    assert np.array_equal(part, orc.hm_partition(g, [2, 4, 3], 0.03, 1, False))


def test_gpu_heipromap_cli_im(tmp_path):
    g = heipa.gen_rgg(20000, 10.0, 8, 10)
    path = str(tmp_path / "rgg.graph")
    heipa.write_metis(path, g)
    out = str(tmp_path / "map.txt")
    stdout = _run([os.path.join(BIN, "GPU-HeiProMap"), "--graph", path, "--hierarchy", "4:8:6", "--distance",
                   "1:10:100", "--imbalance", "0.03", "--config", "IM", "--seed", "1", "--mapping", out,
                   "--verbose-level", "1", "--arena-gb", "1"])
    part = np.loadtxt(out, dtype=np.uint32)
    ref, J, _, _ = orc.im_partition(g, [4, 8, 6], [1, 10, 100], 0.03, 1)
    assert np.array_equal(part, ref)
    assert f"Final comm-cost   : {J}" in stdout
