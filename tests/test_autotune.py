"""Autotuner bookkeeping (CPU, with a stand-in timer) and the struct layout of the C ABI."""

import json
import os
import subprocess
import textwrap

import pytest

from shearwave_amd.autotune import CANDIDATES, Tuner, grid_class

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_grid_class_buckets():
    assert grid_class((64, 64, 64)) == (64, 8)
    assert grid_class((65, 64, 64)) == (128, 8)
    assert grid_class((192, 192, 256)) == (256, 128)
    assert grid_class((100, 20, 70)) == (128, 8)


def test_tuner_picks_fastest_and_caches(tmp_path):
    calls = []

    def fake(shape, xc):
        calls.append((tuple(shape), xc))
        return abs(xc - 48) + 1.0  # fastest at 48

    path = str(tmp_path / "t.json")
    t = Tuner(path=path, timer=fake, arch="gfx950:256")
    assert t.xchunk((128, 96, 128)) == 48
    n = len(calls)
    assert n == len([c for c in CANDIDATES if c <= 128])
    assert t.xchunk((120, 80, 110)) == 48  # same class: no new timing
    assert len(calls) == n
    t2 = Tuner(path=path, timer=fake, arch="gfx950:256")
    assert t2.xchunk((128, 96, 128)) == 48 and len(calls) == n  # reloaded from disk
    table = json.load(open(path))
    assert any(k.startswith("gfx950:256|") for k in table)
    assert t.xchunk((20, 20, 20)) <= 20  # never deeper than the grid


def test_c_struct_layouts_match_ctypes(tmp_path):
    import ctypes as C

    from shearwave_amd import _abi as A

    src = tmp_path / "layout.c"
    src.write_text(textwrap.dedent("""
        #include <stdio.h>
        #include <stddef.h>
        #include "shearwave.h"
        int main(void) {
            printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(sw_fdtd_state), sizeof(sw_fdtd_outputs),
                   sizeof(sw_cg_info), sizeof(sw_batch_solve), offsetof(sw_batch_solve, u),
                   offsetof(sw_batch_solve, env), sizeof(sw_batch_cg), offsetof(sw_batch_cg, part),
                   offsetof(sw_batch_cg, item0), sizeof(sw_static_info), offsetof(sw_static_info, relres),
                   offsetof(sw_static_info, reason));
            return 0;
        }
    """))
    exe = tmp_path / "layout"
    try:
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    except (OSError, subprocess.CalledProcessError):
        pytest.skip("no C compiler")
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(A.FdtdState), C.sizeof(A.FdtdOutputs), C.sizeof(A.CgInfo), C.sizeof(A.BatchSolve),
            A.BatchSolve.u.offset, A.BatchSolve.env.offset, C.sizeof(A.BatchCg), A.BatchCg.part.offset,
            A.BatchCg.item0.offset, C.sizeof(A.StaticInfo), A.StaticInfo.relres.offset, A.StaticInfo.reason.offset]
    assert got == want
