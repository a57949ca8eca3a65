"""The parameter-file driver's SU header scan (index_shot, main.cpp:52-155)
on SU files written here with known headers; no device work."""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ds2d_amd", "bin", "ds2d_lsrtm")


def write_su(path, traces):
    """traces: list of (sx, gx, samples float32); 240-byte SEG-Y trace headers."""
    with open(path, "wb") as f:
        for i, (sx, gx, tr) in enumerate(traces):
            h = bytearray(240)
            struct.pack_into("<i", h, 0, i + 1)
            struct.pack_into("<i", h, 72, sx)
            struct.pack_into("<i", h, 80, gx)
            struct.pack_into("<H", h, 114, tr.size)
            struct.pack_into("<H", h, 116, 800)
            f.write(h)
            f.write(np.asarray(tr, np.float32).tobytes())


@pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built")
def test_index_shot_groups_traces_by_source(tmp_path):
    nt = 16
    rng = np.random.default_rng(0)
    tr = []
    layout = [(1000, [900, 1000, 1100, 1200]), (1500, [1300, 1600]), (2000, [2500, 2600, 2700])]
    for sx, gxs in layout:
        for gx in gxs:
            tr.append((sx, gx, rng.standard_normal(nt)))
    p = tmp_path / "d.su"
    write_su(p, tr)
    out = subprocess.run([EXE, "index", str(p)], capture_output=True, text=True, check=True).stdout.split("\n")
    assert out[0].split() == ["nt", "16", "dt", "0.000800", "shots", "3"]
    rows = [list(map(int, l.split())) for l in out[1:] if l.strip()]
    # sno ntr sx cx_min cx_max pos ; cx_min/max span sources and receivers
    assert rows == [[1, 4, 1000, 900, 1200, 0], [2, 2, 1500, 1300, 1600, 4], [3, 3, 2000, 2000, 2700, 6]]
