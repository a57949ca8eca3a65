"""The chunk planner of the two-step elastic kernel (ds2d_amd/csrc/chunk_plan.hpp),
compiled on the host with g++ through a small C shim: the balanced plan covers
every (window, row block) exactly once with chunks of at least 16 columns,
gives each of the eight XCDs the same number of chunks (a multiple of its CU
count), orders each XCD's chunks widest first, and its makespan is the list
schedule that a brute-force simulation here computes.  On the C3 batch it
finishes sooner than 240-column chunks."""
import ctypes as C
import heapq
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

SHIM = r'''
#include "chunk_plan.hpp"
extern "C" long long plan(const int* nxp, const int* nzp, int n, int rows, int cpx, int W, int balanced,
                          int* out, int cap, int* count) {
    std::vector<int> a(nxp, nxp + n), b(nzp, nzp + n);
    std::vector<ds2d::PlanChunk> c;
    const long long s = balanced ? ds2d::plan_balanced(a, b, rows, cpx, c) : ds2d::plan_fixed(a, b, rows, W, cpx, c);
    *count = (int)c.size();
    for (int i = 0; i < (int)c.size() && i < cap; ++i) {
        out[4 * i] = c[i].slot; out[4 * i + 1] = c[i].x0; out[4 * i + 2] = c[i].x1; out[4 * i + 3] = c[i].z0;
    }
    return s;
}
'''


@pytest.fixture(scope="module")
def planner(tmp_path_factory):
    d = tmp_path_factory.mktemp("plan")
    src = d / "shim.cpp"
    src.write_text(SHIM)
    so = d / "libplan.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "ds2d_amd", "csrc"),
                    str(src), "-o", str(so)], check=True)
    lib = C.CDLL(str(so))
    lib.plan.restype = C.c_longlong

    def run(nxp, nzp, rows, cpx, W=240, balanced=True):
        nxp = np.ascontiguousarray(nxp, np.int32)
        nzp = np.ascontiguousarray(nzp, np.int32)
        cap = 1 << 16
        out = np.zeros(4 * cap, np.int32)
        cnt = C.c_int(0)
        p = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        span = lib.plan(p(nxp), p(nzp), len(nxp), rows, cpx, W, int(balanced), p(out), cap, C.byref(cnt))
        return span, out[:4 * cnt.value].reshape(-1, 4)
    return run


def _makespan(costs, ncu):
    free = [0] * ncu
    end = 0
    for c in costs:
        t = heapq.heappop(free) + c
        heapq.heappush(free, t)
        end = max(end, t)
    return end


def _c3_windows():
    from ds2d_amd import models
    cfg = models.c3_batched_elastic(nshots=128, nt=10)
    return np.array([sh.nx + 2 * cfg["pml"] for sh in cfg["shots"]]), cfg["nz"] + 2 * cfg["pml"]


@pytest.mark.parametrize("cpx", [32, 16, 38])
@pytest.mark.parametrize("case", ["c3", "two_row_blocks", "random"])
def test_balanced_plan_invariants(planner, case, cpx):
    """cpx: CUs per XCD (32 on MI355X; other counts for parts with fewer or more CUs)."""
    rows = 672
    if case == "c3":
        nxp, nz = _c3_windows()
        nzp = np.full(len(nxp), nz)
    elif case == "two_row_blocks":   # C5-like: 1080 rows -> two row blocks per window
        rng = np.random.default_rng(5)
        nxp = rng.integers(480, 1281, 256)
        nzp = np.full(256, 1080)
    else:
        rng = np.random.default_rng(11)
        nxp = rng.integers(200, 2000, 40)
        nzp = rng.integers(100, 1500, 40)
    span, ch = planner(nxp, nzp, rows, cpx)
    assert span > 0
    n = len(ch)
    assert n % (8 * cpx) == 0, n
    K = n // 8
    # exact cover of every (window, row block) by [x0, x1) ranges, widths >= 16
    assert np.all(ch[:, 2] - ch[:, 1] >= 16)
    for s in range(len(nxp)):
        for z0 in range(0, nzp[s], rows):
            m = (ch[:, 0] == s) & (ch[:, 3] == z0)
            iv = sorted(zip(ch[m, 1], ch[m, 2]))
            assert iv[0][0] == 0 and iv[-1][1] == nxp[s]
            assert all(a[1] == b[0] for a, b in zip(iv, iv[1:]))
    assert set(ch[:, 3]) <= set(range(0, int(nzp.max()), rows))
    # each XCD: widest first, and the reported span is the simulated list schedule
    spans = []
    for x in range(8):
        w = ch[x * K:(x + 1) * K, 2] - ch[x * K:(x + 1) * K, 1]
        assert np.all(np.diff(w) <= 0)
        spans.append(_makespan(list(w + 24), cpx))
        # a window's row block lives on one XCD
        keys = {(a, b) for a, b in zip(ch[x * K:(x + 1) * K, 0], ch[x * K:(x + 1) * K, 3])}
        for y in range(8):
            if y != x:
                other = {(a, b) for a, b in zip(ch[y * K:(y + 1) * K, 0], ch[y * K:(y + 1) * K, 3])}
                assert not keys & other
    assert span == max(spans)


def test_fixed_plan_matches_the_launch_layout(planner):
    nxp, nz = _c3_windows()
    nzp = np.full(len(nxp), nz)
    span, ch = planner(nxp, nzp, 672, 32, W=240, balanced=False)
    assert len(ch) == sum(-(-int(v) // 240) for v in nxp)
    per = -(-len(ch) // 8)
    w = ch[:, 2] - ch[:, 1]
    assert span == max(_makespan(list(w[x * per:(x + 1) * per] + 24), 32) for x in range(8))


def test_balanced_plan_beats_fixed_width_on_c3(planner):
    """Measured on C3 (two steps per launch, 704 lanes): 240-column chunks end
    at 1487 us per launch against a mean CU busy time of 1306 us; the model
    (w + 24 column-times per chunk) puts the plan 9 % below fixed width."""
    nxp, nz = _c3_windows()
    nzp = np.full(len(nxp), nz)
    fixed, _ = planner(nxp, nzp, 672, 32, W=240, balanced=False)
    bal, ch = planner(nxp, nzp, 672, 32)
    total = int(sum(nxp)) + 24 * len(ch)
    print(f"C3: fixed 240 -> {fixed}, balanced -> {bal} ({len(ch)} chunks), mean per CU {total / 256:.0f}")
    assert bal < 0.93 * fixed
    assert bal <= 1.04 * total / 256


def test_small_batches_fall_back_to_fixed_width(planner):
    span, ch = planner([300, 280], [120, 120], 672, 32)
    assert span == -1 and len(ch) == 0
