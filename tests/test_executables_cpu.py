"""The drop-in programs' paths that end before any GPU work, on the CPU.

The option table and the template writer (inputParser.cpp:45-88; the option
tables of langevinDynamicsSingleBarrier.cpp:65-108 and
langevinDynamicsDoubleBarrier.cpp:80-85): `-o file` rewrites an existing input
file from its own values with the help comments, so the reference's sample
inputs, which that writer produced, come back byte for byte (the old file is
kept as `#name_bk.1#`, MOLUTILITY.cpp:6-22).  The runner's early exits
(Single.h:3221-3224 numberOfFrame < 1000, :3327-3332 fewer than 20 lags) end
the run with status 0 and the reference's message.
"""
import os
import re
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "langevindynamicsforpcp_amd", "bin")
GOLD = os.path.join(os.path.dirname(__file__), "golden")
SINGLE = os.path.join(BIN, "langevinDynamicsSingleBarrier")
DOUBLE = os.path.join(BIN, "langevinDynamicsDoubleBarrier")


@pytest.fixture(autouse=True)
def built():
    if not (os.path.exists(SINGLE) and os.path.exists(DOUBLE)):
        pytest.fail("executables not built: run python -m langevindynamicsforpcp_amd.build")


def run(exe, args, cwd):
    return subprocess.run([exe, *args], cwd=cwd, capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("exe,name", [(SINGLE, "quick_test.inp"), (DOUBLE, "cuidb_double_2.0.inp")])
def test_template_rewrites_sample_input_byte_for_byte(tmp_path, exe, name):
    src = os.path.join(GOLD, name)
    shutil.copyfile(src, tmp_path / "x.inp")
    r = run(exe, ["-o", "x.inp"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "x.inp").read_bytes() == open(src, "rb").read()
    assert (tmp_path / "#x.inp_bk.1#").read_bytes() == open(src, "rb").read()


def test_template_from_scratch_lists_every_key(tmp_path):
    r = run(DOUBLE, ["-o", "new.inp"], tmp_path)
    assert r.returncode == 0, r.stderr
    text = (tmp_path / "new.inp").read_text()
    keys = [ln.split()[0] for ln in text.splitlines() if ln and not ln.startswith("#")]
    assert keys[:3] == ["OUTPUTBASE", "OVERDAMPED", "X0"]
    for k in ("BARRIERHEIGHTS0", "BARRIERHEIGHTS1", "BARRIERPOSITION0", "BARRIERPOSITION1", "FPTCG", "RANDOMSEED"):
        assert k in keys
    assert "BARRIERHEIGHTS" not in keys and "BARRIERHALFWIDTH" not in keys
    # every key line follows its "# [Kind] help" comment
    lines = text.splitlines()
    for i, ln in enumerate(lines):
        if ln and not ln.startswith("#"):
            assert re.match(r"# \[(Input|Output|Parameter)\] ", lines[i - 1]), ln


def test_print_values_does_not_run(tmp_path):
    shutil.copyfile(os.path.join(GOLD, "quick_test.inp"), tmp_path / "x.inp")
    r = run(SINGLE, ["-i", "x.inp", "-p"], tmp_path)
    assert r.returncode == 0
    assert re.search(r"BARRIERHEIGHTS\s+0 7\.20 2\.00", r.stderr)
    assert not any(f.endswith(".dat") for f in os.listdir(tmp_path))


def test_input_and_output_together_is_an_error(tmp_path):
    shutil.copyfile(os.path.join(GOLD, "quick_test.inp"), tmp_path / "x.inp")
    r = run(SINGLE, ["-i", "x.inp", "-o", "y.inp"], tmp_path)
    assert "only one of -i and -o" in r.stderr and not (tmp_path / "y.inp").exists()


def test_too_few_frames_ends_the_run(tmp_path):
    text = open(os.path.join(GOLD, "quick_test.inp")).read()
    text = re.sub(r"(?m)^FREQSAVE .*$", "FREQSAVE 10000000", text)   # 3e9 / 1e7 + 1 = 301 frames
    (tmp_path / "x.inp").write_text(text)
    r = run(SINGLE, ["-i", "x.inp"], tmp_path)
    assert r.returncode == 0 and "Error: numberOfFrame < 1000" in r.stderr
    assert not any(f.endswith(".dat") for f in os.listdir(tmp_path))


def test_too_few_lags_ends_the_run(tmp_path):
    text = open(os.path.join(GOLD, "quick_test.inp")).read()
    text = re.sub(r"(?m)^NUMBEROFSTEPS .*$", "NUMBEROFSTEPS 300300000", text)   # 1002 frames: lags span < 1 decade
    text = re.sub(r"(?m)^NUMBEROFLOGSAMPLES .*$", "NUMBEROFLOGSAMPLES 2", text)
    (tmp_path / "x.inp").write_text(text)
    r = run(SINGLE, ["-i", "x.inp"], tmp_path)
    assert r.returncode == 0 and "Error: interval is less than 20" in r.stderr
    assert not any(f.endswith(".dat") for f in os.listdir(tmp_path))
