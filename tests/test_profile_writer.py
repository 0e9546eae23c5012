"""The .npz profile writer of the host layer (the Archiver stand-in),
compiled with the host compiler alone and read back by numpy.load (CPU)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "dspsr-kanyixia_amd", "host")

MAIN = r"""
#include "dsp/ProfileWriter.h"
#include <vector>
int main(int argc, char** argv)
{
  std::vector<float> p(2 * 3 * 4 * 5);
  for (size_t i = 0; i < p.size(); i++) p[i] = float(i) * 0.25f - 7.0f;
  std::vector<uint32_t> h(2 * 5);
  for (size_t i = 0; i < h.size(); i++) h[i] = uint32_t(i * 1000003u);
  std::vector<float> one(1, 3.5f);
  dsp::ProfileWriter w;
  w.add("profiles", p.data(), {2, 3, 4, 5});
  w.add("hits", h.data(), {2, 5});
  w.add("vector", one.data(), {1});
  w.add("folding_period", 0.253);
  w.write(argv[1]);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_npz_round_trip(tmp_path):
    src = tmp_path / "main.C"
    src.write_text(MAIN)
    exe = tmp_path / "w"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", HOST, str(src), os.path.join(HOST, "ProfileWriter.C"), "-o",
                    str(exe)], check=True)
    out = tmp_path / "fold.npz"
    subprocess.run([str(exe), str(out)], check=True)
    z = np.load(out, allow_pickle=False)
    assert sorted(z.files) == ["folding_period", "hits", "profiles", "vector"]
    p = z["profiles"]
    assert p.dtype == np.float32 and p.shape == (2, 3, 4, 5)
    assert np.array_equal(p.ravel(), np.arange(120, dtype=np.float32) * np.float32(0.25) - np.float32(7.0))
    h = z["hits"]
    assert h.dtype == np.uint32 and np.array_equal(h.ravel(), (np.arange(10, dtype=np.uint64) * 1000003).astype(np.uint32))
    assert z["vector"].shape == (1,) and z["vector"][0] == 3.5
    assert z["folding_period"].shape == () and float(z["folding_period"]) == 0.253
