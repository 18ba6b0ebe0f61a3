"""The C-ABI library builds for gfx950, loads, and exports every entry point
include/ldpcp.h declares; the ctypes structures match the C layout."""
import os
import re
import subprocess

import pytest

from langevindynamicsforpcp_amd import _native as N

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "ldpcp.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)  # header-only helpers
    return sorted(set(re.findall(r"\b(ldpcp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(N.LIB_PATH):
        pytest.fail("libldpcp.so not built: run python -m langevindynamicsforpcp_amd.build")
    out = subprocess.check_output(["nm", "-D", "--defined-only", N.LIB_PATH], text=True)
    exported = set(re.findall(r"\sT\s(ldpcp_\w+)", out))
    declared = declared_functions()
    assert declared, "no declarations parsed"
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    assert sorted(N.symbols()) == declared


def test_library_loads_without_gpu():
    lib = N.load()
    assert lib.ldpcp_last_error() is not None


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include "ldpcp.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(ldpcp_run_params), '
                   'sizeof(ldpcp_lane), sizeof(ldpcp_sweep_set), sizeof(ldpcp_set_stats), '
                   'offsetof(ldpcp_lane, wt_tick_prev), offsetof(ldpcp_run_params, barrier_window), '
                   'sizeof(ldpcp_engine_config), sizeof(ldpcp_server_config), offsetof(ldpcp_set_stats, hist), '
                   'offsetof(ldpcp_server_config, engine), sizeof(ldpcp_checkpoint_header), '
                   'offsetof(ldpcp_checkpoint_header, timestep));return 0;}\n')
    exe = tmp_path / "sizes"
    subprocess.check_call(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)])
    got = [int(v) for v in subprocess.check_output([str(exe)], text=True).split()]
    import ctypes as C
    want = [C.sizeof(N.RunParams), C.sizeof(N.LaneResult), C.sizeof(N.SweepSet), C.sizeof(N.SetStats),
            N.LaneResult.wt_tick_prev.offset, N.RunParams.barrier_window.offset, C.sizeof(N.EngineConfig),
            C.sizeof(N.ServerConfig), N.SetStats.hist.offset, N.ServerConfig.engine.offset,
            C.sizeof(N.CheckpointHeader), N.CheckpointHeader.timestep.offset]
    assert got == want
