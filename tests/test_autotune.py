"""Autotuner bookkeeping (CPU, with a stand-in timer) and the struct layout of the C ABI."""

import json
import os
import subprocess
import textwrap

import pytest

from shearwave_amd.autotune import BATCH_ROWS, CAPS, PASS_ROWS, PASS_SPLITS, Tuner, balanced_chunk, batch_candidates, \
    pass_candidates, tune_mix_shapes

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


def test_tuner_picks_fastest_batch_shape_and_caches(tmp_path):
    calls = []

    def fake(cands):
        calls.append(tuple(cands))
        return {(r, c): abs(c - 128) + abs(r - 8) * 100 + 1.0 for r, c in cands}  # fastest at (8, 128)

    path = str(tmp_path / "t.json")
    t = Tuner(path=path, timer=fake, arch="gfx950:256")
    assert t.batch_shape() == (8, 128) and calls == [tuple(batch_candidates())]
    assert t.cap() == 128 and t.tile_y() == 8
    assert {r for r, _ in calls[0]} == set(BATCH_ROWS) and {c for _, c in calls[0]} == set(CAPS)
    assert t.xchunk((200, 64, 64)) == 100 and t.xchunk((100, 64, 64)) == 100
    assert len(calls) == 1  # measured once
    t2 = Tuner(path=path, timer=fake, arch="gfx950:256")
    assert t2.batch_shape() == (8, 128) and len(calls) == 1  # reloaded from disk
    table = json.load(open(path))
    e = table["gfx950:256|batch_shape"]
    assert e["cap"] == 128 and e["tile_y"] == 8 and len(e["ps_per_voxel_step"]) == len(BATCH_ROWS) * len(CAPS)
    t3 = Tuner(path=path, timer=fake, arch="gfx950:304")  # another device is tuned anew
    assert t3.batch_shape() == (8, 128) and len(calls) == 2


def test_tuner_can_pick_other_tile_rows(tmp_path):
    t = Tuner(path=str(tmp_path / "t.json"), timer=lambda cands: {c: (1.0 if c == (12, 64) else 2.0) for c in cands},
              arch="x")
    assert t.batch_shape() == (12, 64)


class _FakeSolve:
    def __init__(self, shape, dtype="float32", viscous=True):
        self.shape, self.dtype, self.e = shape, dtype, (object() if viscous else None)


def test_pass_candidates():
    c = pass_candidates("float32", (256, 256, 256))
    assert {r for r, _ in c} == set(PASS_ROWS["float32"]) and len(c) == len(set(c))
    assert sorted({xc for _, xc in c}) == [0, 16, 22, 32, 43, 64, 86, 128, 256]  # one wave, and equal splits
    for _, xc in c:  # every depth but 0 cuts the axis into equal chunks
        assert xc == 0 or any(-(-256 // k) == xc for k in PASS_SPLITS)
    c = pass_candidates("float32", (40, 64, 64))  # splits below PASS_MIN_XCHUNK planes are left out
    assert sorted({xc for _, xc in c}) == [0, 10, 14, 20, 40]
    assert sorted({xc for _, xc in pass_candidates("float32", (6, 64, 64))}) == [0, 6]
    assert {r for r, _ in pass_candidates("float64", (64, 64, 64))} == set(PASS_ROWS["float64"])


def test_pass_shape_per_grid_class_and_cached(tmp_path):
    calls = []

    def fake(solve, cands):
        calls.append((solve.shape, tuple(cands)))
        want = (12, 64) if solve.shape[0] == 256 else (8, 0)  # both among the candidates of these shapes
        return {c: (1.0 if c == want else 2.0 + c[1]) for c in cands}

    path = str(tmp_path / "t.json")
    t = Tuner(path=path, pass_timer=fake, arch="gfx950:256")
    assert t.pass_shape(_FakeSolve((256, 256, 256))) == (12, 64)
    assert t.pass_shape(_FakeSolve((128, 96, 160))) == (8, 0)
    assert t.pass_shape(_FakeSolve((256, 256, 256))) == (12, 64) and len(calls) == 2
    assert calls[0][1] == tuple(pass_candidates("float32", (256, 256, 256)))
    # storage type and viscosity are separate classes
    t.pass_shape(_FakeSolve((256, 256, 256), viscous=False))
    t.pass_shape(_FakeSolve((256, 256, 256), dtype="float64"))
    assert len(calls) == 4 and calls[3][1] == tuple(pass_candidates("float64", (256, 256, 256)))
    t2 = Tuner(path=path, pass_timer=fake, arch="gfx950:256")
    assert t2.pass_shape(_FakeSolve((128, 96, 160))) == (8, 0) and len(calls) == 4  # from disk
    e = t2.entry(t2.pass_key("float32", True, (256, 256, 256)))
    assert e["tile_y"] == 12 and e["x_chunk"] == 64 and "12x64" in e["ps_per_voxel_step"]


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
