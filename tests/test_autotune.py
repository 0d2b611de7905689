"""Autotuner bookkeeping (CPU, with a stand-in timer) and the struct layout of the C ABI."""

import json
import os
import subprocess
import textwrap

import pytest

from shearwave_amd.autotune import CAPS, Tuner, balanced_chunk, tune_mix_shapes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_balanced_chunk():
    assert balanced_chunk(96, 96) == 96
    assert balanced_chunk(97, 96) == 49  # two equal chunks, not 96 + 1
    assert balanced_chunk(150, 96) == 75
    assert balanced_chunk(256, 96) == 86  # three chunks
    assert balanced_chunk(40, 96) == 40  # never deeper than the grid
    for nx in range(1, 300):
        for cap in CAPS:
            c = balanced_chunk(nx, cap)
            n = -(-nx // c)
            assert c <= max(cap, 1) and n == -(-nx // cap) and c * (n - 1) < nx  # fewest chunks, none empty


def test_tune_mix_is_fixed():
    shapes = tune_mix_shapes()
    assert shapes == tune_mix_shapes() and len(shapes) == 24
    assert all(64 <= v <= 256 and v % 8 == 0 for sh in shapes for v in sh)
    assert len(set(shapes)) > 20  # uneven grids


def test_tuner_picks_fastest_cap_and_caches(tmp_path):
    calls = []

    def fake(caps):
        calls.append(tuple(caps))
        return {c: abs(c - 128) + 1.0 for c in caps}  # fastest at 128

    path = str(tmp_path / "t.json")
    t = Tuner(path=path, timer=fake, arch="gfx950:256")
    assert t.cap() == 128 and calls == [CAPS]
    assert t.xchunk((200, 64, 64)) == 100 and t.xchunk((100, 64, 64)) == 100
    assert len(calls) == 1  # measured once
    t2 = Tuner(path=path, timer=fake, arch="gfx950:256")
    assert t2.cap() == 128 and len(calls) == 1  # reloaded from disk
    table = json.load(open(path))
    assert table["gfx950:256|batch_cap"]["cap"] == 128
    t3 = Tuner(path=path, timer=fake, arch="gfx950:304")  # another device is tuned anew
    assert t3.cap() == 128 and len(calls) == 2


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
