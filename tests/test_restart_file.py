"""Restart files of the drop-in runner, on the CPU (host code only).

The settings header must decode with the decoder that reads the reference's
own restart files (tests/golden/restart_decode.py, which consumed the shipped
quick-test files byte for byte), so field order and encoding are the
reference's; the body is the library's checkpoint image between the
reference's markers 9998 and 9990.  The read side must apply the reference's
consistency checks with its messages.
"""
import os
import struct
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(REPO, "langevindynamicsforpcp_amd", "host")
sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import restart_decode  # noqa: E402

HARNESS = r"""
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "main_common.h"
#include "restart_file.h"
using namespace ldpcp_host;
// import <inp> <restart base> <rank> <W> <image out>: the configuration of the
// input file, the restart file of (rank, trial 0) as a checkpoint image
static int import_mode(char** argv)
{
    InputParser ip;
    add_options(ip, 0);
    char* args[] = {argv[0], (char*)"-i", argv[2]};
    if (!ip.Parse(3, args, false)) return 3;
    RunnerConfig c = read_config(ip, 0);
    c.input_restart = argv[3];
    ImageSpec spec;
    spec.seed = 77;
    spec.traj_offset = 0;
    spec.barrier_window = 4096;
    spec.n_fields = c.write_x ? 7 : 1;
    spec.tick_capacity = 1 << 16;
    std::vector<char> image;
    uint64_t n = 0;
    if (!read_restart_file(c, atoi(argv[5]), atoi(argv[4]), 0, spec, image, &n)) return 2;
    FILE* f = fopen(argv[6], "wb");
    fwrite(image.data(), 1, image.size(), f);
    fclose(f);
    printf("%llu\n", (unsigned long long)n);
    return 0;
}
int main(int argc, char** argv)
{
    if (!strcmp(argv[1], "import")) return import_mode(argv);
    RunnerConfig c;
    c.model = atoi(argv[2]);
    c.x0 = -1.05; c.v0 = 0.5; c.gamma = 2.0; c.kb = 1.5; c.temperature = 0.75; c.mass = 3.0;
    c.timestep = 1e-3; c.stat_freq = 1; c.save_freq = 10; c.waiting = true; c.write_x = true;
    c.randomize_v = true; c.randomize_state = false; c.same_initial_barrier = true;
    c.periodic_length = 2.1;
    c.height[0][0] = 3.0; c.height[0][1] = 1.0; c.height[1][0] = 2.0; c.height[1][1] = 0.5;
    c.half_width[0] = 0.7; c.half_width[1] = 0.65; c.position[0] = -1.42; c.position[1] = -0.56;
    c.initial_state = 1.0; c.tau = 1e-3;
    c.output_restart = c.input_restart = "rs";
    if (argc > 3) c.timestep = atof(argv[3]);
    std::vector<char> image(1000);
    for (size_t i = 0; i < image.size(); ++i) image[i] = (char)(i * 7 + 3);
    if (!strcmp(argv[1], "write")) return write_restart_file(c, 4, 1, 2, 123456, image) ? 0 : 1;
    std::vector<char> got;
    uint64_t n = 0;
    if (!read_restart_file(c, 4, 1, 2, ImageSpec(), got, &n)) return 2;
    printf("%llu %d\n", (unsigned long long)n, got == image ? 1 : 0);
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("restart")
    (d / "h.cpp").write_text(HARNESS)
    exe = d / "h"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(REPO, "include"), "-I", HOST,
                           str(d / "h.cpp"), os.path.join(HOST, "restart_file.cpp"),
                           os.path.join(HOST, "input_parser.cpp"), os.path.join(HOST, "text_util.cpp"),
                           "-o", str(exe)])
    return exe


@pytest.mark.parametrize("model", [0, 1])
def test_restart_header_decodes_as_the_reference_layout(harness, tmp_path, model):
    subprocess.check_call([str(harness), "write", str(model)], cwd=tmp_path)
    path = tmp_path / "rs_00001_00002.restart"      # <OUTPUTRESTART>_%05d_%05d.restart (rank, trial)
    buf = path.read_bytes()
    r = restart_decode._Reader(buf)
    h = restart_decode.read_header(r, model)
    assert h["signature"] == 0x4C445043 and h["simd_width"] == 4 and h["rank"] == 1
    assert (h["x0"], h["v0"], h["gamma"], h["kb"], h["temperature"], h["mass"], h["timestep"]) == \
        (-1.05, 0.5, 2.0, 1.5, 0.75, 3.0, 1e-3)
    assert h["nsteps_total"] == 123456 and h["stat_freq"] == 1 and h["save_freq"] == 10 and h["waiting_time"] == 1
    assert h["save_flags"] == [1, 1, 1, 1] and h["save_flags2"] == [0, 1, 1]
    assert (h["randomize_v"], h["randomize_eta"], h["randomize_state"], h["same_initial_barrier"]) == (1, 1, 0, 1)
    assert (h["q"], h["tau"], h["eta0"]) == (0.2, 1e-3, 0.0)
    if model == 0:
        assert (h["periodic_length"], h["h0"], h["h1"], h["initial_state"]) == (2.1, 3.0, 1.0, 1.0)
    else:
        assert (h["h00"], h["h01"], h["h10"], h["h11"]) == (3.0, 1.0, 2.0, 0.5)
        assert (h["position0"], h["position1"], h["initial_state"]) == (-1.42, -0.56, 1.0)
        # apex of the inner slopes (Double.h:1144): x where the two slopes meet
        x1h, x2l, y1, y2 = -1.42 + 0.7, -0.56 - 0.65, 3.0, 2.0
        s1, s2 = -y1 / (x1h - (-1.42)), y2 / (-0.56 - x2l)   # falling, rising slope
        assert abs(h["xc"] - (x1h * s1 - x2l * s2) / (s1 - s2)) < 1e-12
    assert h["rng"] == [(123456, 0), (123456, 0), (0, 0)]
    assert r.i32() == 9998
    n = r.u64()
    assert n == 1000 and buf[r.pos:r.pos + n] == bytes((i * 7 + 3) & 0xFF for i in range(1000))
    r.pos += n
    assert r.i32() == 9990 and r.pos == len(buf)


def test_restart_read_round_trip_and_checks(harness, tmp_path):
    subprocess.check_call([str(harness), "write", "0"], cwd=tmp_path)
    out = subprocess.run([str(harness), "read", "0"], cwd=tmp_path, capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.split() == ["123456", "1"]
    # another time step: the reference's message, and no image
    out = subprocess.run([str(harness), "read", "0", "0.002"], cwd=tmp_path, capture_output=True, text=True)
    assert out.returncode == 2 and "Error: inconsistent time step0" in out.stderr
    # another signature is refused at the first check
    path = tmp_path / "rs_00001_00002.restart"
    buf = bytearray(path.read_bytes())
    buf[0:4] = struct.pack("<i", 2346)
    path.write_bytes(bytes(buf))
    out = subprocess.run([str(harness), "read", "0"], cwd=tmp_path, capture_output=True, text=True)
    assert out.returncode == 2 and "Error: inconsistent signature" in out.stderr
    # this program's body under the reference program's signature does not parse as a reference body
    buf[0:4] = struct.pack("<i", 2345)
    path.write_bytes(bytes(buf))
    out = subprocess.run([str(harness), "read", "0"], cwd=tmp_path, capture_output=True, text=True)
    assert out.returncode == 2 and "Error: " in out.stderr
    # a truncated file is refused
    buf[0:4] = struct.pack("<i", 0x4C445043)
    path.write_bytes(bytes(buf[:-10]))
    out = subprocess.run([str(harness), "read", "0"], cwd=tmp_path, capture_output=True, text=True)
    assert out.returncode == 2 and "corrupted" in out.stderr
    # a missing file
    empty = tmp_path / "empty"
    empty.mkdir()
    out = subprocess.run([str(harness), "read", "0"], cwd=empty, capture_output=True, text=True)
    assert out.returncode == 2 and "can't open restart file (rs_00001_00002.restart)" in out.stderr


QUICK_JOB = "/root/reference/sample_input/quick_test/executed_job"
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def parse_image(buf):
    """Checkpoint image -> header, lanes, window base, window, frames, ticks."""
    import ctypes as C
    import numpy as np
    from langevindynamicsforpcp_amd import _native as N
    hd = N.CheckpointHeader.from_buffer_copy(buf)
    n, wb, nf, cap = hd.n_traj, hd.barrier_window, hd.n_frames, hd.tick_capacity
    o = C.sizeof(N.CheckpointHeader)
    lanes = (N.LaneResult * n).from_buffer_copy(buf[o:o + n * C.sizeof(N.LaneResult)])
    o += n * C.sizeof(N.LaneResult)
    wbase = np.frombuffer(buf, np.int32, n, o)
    o += (n * 4 + 7) // 8 * 8
    win = np.frombuffer(buf, np.uint64, n * 2 * wb, o).reshape(n, 2, wb)
    o += n * 2 * wb * 8
    frames = np.frombuffer(buf, np.float64, hd.n_fields * nf * n, o).reshape(hd.n_fields, nf, n)
    o += hd.n_fields * nf * n * 8
    ticks = np.frombuffer(buf, np.float64, n * cap, o).reshape(n, cap)
    assert o + n * cap * 8 == len(buf) == hd.bytes
    return hd, lanes, wbase, win, frames, ticks


@pytest.mark.skipif(not os.path.isdir(QUICK_JOB), reason="the reference's recorded quick-test job is not on this machine")
def test_reference_restart_files_import_to_the_recorded_state(harness, tmp_path):
    """The reference's own quick-test restart files (4 ranks x 4 lanes, 3e9
    steps) read into checkpoint images: every lane's state, statistics sums and
    counts, frames, barrier table and R_prev equal what the fixtures recorded
    from the same files, and the per-lane means equal the reference's .dat
    outputs bit for bit (the device's running sums are the reference's
    STATISTICS::AVERAGE)."""
    import json
    import numpy as np
    import restart_decode as R
    st = json.load(open(os.path.join(GOLDEN, "quick_test_state.json")))
    frames_x = np.load(os.path.join(GOLDEN, "quick_test_frames.npz"))["x"]
    out = st["outputs"]
    base = os.path.join(QUICK_JOB, "m1.0_t1.00e-05_restart_0")
    for rank in range(4):
        # the encoder writes the decoded file back byte for byte
        src = f"{base}_{rank:05d}_00000.restart"
        R.encode(R.decode(src), str(tmp_path / "again.restart"))
        assert (tmp_path / "again.restart").read_bytes() == open(src, "rb").read()
        img = tmp_path / f"img{rank}"
        r = subprocess.run([str(harness), "import", os.path.join(GOLDEN, "quick_test.inp"), base, str(rank), "4",
                            str(img)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        assert int(r.stdout) == 3000000000
        hd, lanes, wbase, win, frames, ticks = parse_image(img.read_bytes())
        assert (hd.steps_done, hd.n_frames, hd.n_traj, hd.seed, hd.n_fields) == (3000000000, 10001, 4, 77, 7)
        rk = st["ranks"][rank]
        for i in range(4):
            L, want = lanes[i], st["lanes"][rank * 4 + i]
            g = rank * 4 + i
            assert (L.x, L.state, L.well, L.periodic) == (want["x"], want["state"], want["well"], want["periodic"])
            assert list(L.state_count) == want["state_count"]
            assert (L.v0_instance, L.state0_instance) == (want["v0_instance"], want["state0_instance"])
            assert (L.fpt_n, L.fpt2_n, L.tpt_n, L.tpt2_n) == (want["fpt_n"], want["fpt2_n"], want["tpt_n"],
                                                             want["tpt2_n"])
            assert L.fpt_sum == pytest.approx(want["fpt_sum"], rel=1e-12)  # the fixture summed pairwise
            assert (L.fpt_last_tick, L.fpt_to_update) == (want["fpt_last_tick"], want["fpt_to_update"])
            assert list(L.fpt2_idx) == want["fpt2_index"]
            assert L.fpt_sum / L.fpt_n == out["firstPassageTime.dat"][g]
            assert L.fpt2_sum / L.fpt2_n == out["firstPassageTime2.dat"][g]
            assert L.tpt_sum / L.tpt_n == out["transitionPathTime.dat"][g]
            assert L.tpt2_sum / L.tpt2_n == out["transitionPathTime2.dat"][g]
            for j, size in enumerate(want["cg_sizes"]):
                assert L.cg_n[j] == want["cg_n"][j]
                name = f"firstPassageTimeCG2_{j:04d}_{size:04d}.dat"
                assert (L.cg_sum[j] / L.cg_n[j] if L.cg_n[j] else 0.0) == out[name][g], name
            assert L.wt_ticks == want["ticks_n"]
            np.testing.assert_array_equal(ticks[i, :20], want["ticks_first"])
            assert L.r_prev == rk["r_prev"][i]
            np.testing.assert_array_equal(frames[0, :, i], frames_x[g])
            # barrier table: every recorded entry sits in the lane's window with its state and step
            for k in range(rk["barrier_n"]):
                e = int(win[i, 0, rk["barrier_min_id"] + k - wbase[i]])
                assert e >> 63 == 1 and (e >> 62) & 1 == rk["barrier_state"][k][i]
                assert e & ((1 << 62) - 1) == rk["barrier_step"][k][i]


def test_host_restart_code_is_clean_under_sanitizers(tmp_path):
    """The same harness built with AddressSanitizer and UndefinedBehaviorSanitizer
    (host code only): writing, reading, rejecting a truncated file and, where
    the reference's recorded files are present, importing one of them."""
    (tmp_path / "h.cpp").write_text(HARNESS)
    exe = tmp_path / "h_san"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=undefined", "-I", os.path.join(REPO, "include"), "-I", HOST,
                           str(tmp_path / "h.cpp"), os.path.join(HOST, "restart_file.cpp"),
                           os.path.join(HOST, "input_parser.cpp"), os.path.join(HOST, "text_util.cpp"),
                           "-o", str(exe)])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")

    def go(*args, rc=0):
        r = subprocess.run([str(exe), *args], cwd=tmp_path, capture_output=True, text=True, env=env)
        assert r.returncode == rc and "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
        return r

    go("write", "0")
    assert go("read", "0").stdout.split() == ["123456", "1"]
    path = tmp_path / "rs_00001_00002.restart"
    path.write_bytes(path.read_bytes()[:-10])
    go("read", "0", rc=2)
    if os.path.isdir(QUICK_JOB):
        go("import", os.path.join(GOLDEN, "quick_test.inp"), os.path.join(QUICK_JOB, "m1.0_t1.00e-05_restart_0"),
           "2", "4", str(tmp_path / "img"))
