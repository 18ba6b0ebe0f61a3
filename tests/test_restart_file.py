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
#include "ldpcp.h"
#include "main_common.h"
#include "restart_file.h"
using namespace ldpcp_host;
static RunnerConfig config_of(char* prog, char* inp, int model)
{
    InputParser ip;
    add_options(ip, model);
    char* args[] = {prog, (char*)"-i", inp};
    if (!ip.Parse(3, args, false)) exit(3);
    return read_config(ip, model);
}
// import <inp> <restart base> <rank> <W> <image out> [model [event capacity]]: the
// configuration of the input file, the restart file of (rank, trial 0) as a
// checkpoint image
static int import_mode(int argc, char** argv)
{
    const int model = argc > 7 ? atoi(argv[7]) : 0;
    RunnerConfig c = config_of(argv[0], argv[2], model);
    c.input_restart = argv[3];
    ImageSpec spec;
    spec.seed = 77;
    spec.traj_offset = 0;
    spec.barrier_window = 4096;
    spec.n_fields = c.write_x ? 7 : 1;
    spec.tick_capacity = 1 << 16;
    if (argc > 8) {
        spec.event_capacity = strtoull(argv[8], nullptr, 10);
        spec.n_lists = LDPCP_LIST_CG0 + (c.fptcg ? (uint32_t)c.cg_size.size() : 0u);
    }
    std::vector<char> image;
    uint64_t n = 0;
    if (!read_restart_file(c, atoi(argv[5]), atoi(argv[4]), 0, spec, image, &n)) return 2;
    FILE* f = fopen(argv[6], "wb");
    fwrite(image.data(), 1, image.size(), f);
    fclose(f);
    printf("%llu\n", (unsigned long long)n);
    return 0;
}
// export <inp> <model> <image> <tables> <restart base> <rank> <W>: the image
// written in the reference's layout (trial 0) with the barrier tables of the
// file <tables> (per array: int32 lo, int32 n, states [W][n] int32, steps
// [W][n] uint64)
static int export_mode(char** argv)
{
    const int model = atoi(argv[3]);
    RunnerConfig c = config_of(argv[0], argv[2], model);
    c.output_restart = argv[6];
    const int W = atoi(argv[8]);
    std::vector<char> image;
    FILE* f = fopen(argv[4], "rb");
    for (int ch; (ch = fgetc(f)) != EOF;) image.push_back((char)ch);
    fclose(f);
    if (!reference_layout_possible(c, image)) return 4;
    std::vector<BarrierTable> tables(model ? 2 : 1);
    f = fopen(argv[5], "rb");
    for (BarrierTable& T : tables) {
        if (fread(&T.lo, 4, 1, f) != 1 || fread(&T.n, 4, 1, f) != 1) return 5;
        T.states.resize((size_t)W * T.n);
        T.steps.resize((size_t)W * T.n);
        if (fread(T.states.data(), 4, T.states.size(), f) != T.states.size()) return 5;
        if (fread(T.steps.data(), 8, T.steps.size(), f) != T.steps.size()) return 5;
    }
    fclose(f);
    ldpcp_checkpoint_header hd;
    memcpy(&hd, image.data(), sizeof(hd));
    return write_reference_restart_file(c, W, atoi(argv[7]), 0, hd.steps_done, image, tables) ? 0 : 1;
}
int main(int argc, char** argv)
{
    if (!strcmp(argv[1], "import")) return import_mode(argc, argv);
    if (!strcmp(argv[1], "export")) return export_mode(argv);
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
    o += n * cap * 8
    if hd.n_lists:   # the event lists, trajectory after trajectory, list after list
        counts = [min(_list_count(lanes[t], k), hd.event_capacity) for t in range(n) for k in range(hd.n_lists)]
        o += 8 * sum(counts)
    assert o == len(buf) == hd.bytes
    return hd, lanes, wbase, win, frames, ticks


def _list_count(lane, k):
    """ldpcp_list_count: FPT1 and its two barrier logs share fpt_n."""
    return {0: lane.fpt_n, 1: lane.fpt2_n, 2: lane.tpt_n, 3: lane.tpt2_n, 4: lane.fpt_n,
            5: lane.fpt_n}.get(k, lane.cg_n[k - 6] if k >= 6 else 0)


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
    # the double barrier's reference layout: import, export, and the exported file read back
    import numpy as np
    import restart_decode as R
    rng = np.random.default_rng(4)
    W, n = 4, 301
    xc = -1.0
    lanes = [_double_lane(rng, i, 3 * 10**9, 11, [2, 3, 4, 5, 10, 100, 1000], xc) for i in range(W)]
    tab = [(rng.integers(0, 2, (n, W)).astype(np.int32), rng.integers(0, 10**6, (n, W)).astype(np.uint64))
           for _ in range(2)]
    h = {"signature": 2345, "simd_width": W, "x0": -2.9, "v0": 1.0, "gamma": 1.0, "kb": 1.0, "temperature": 1.0,
         "mass": 1.0, "timestep": 1e-4, "nsteps_total": 3 * 10**9, "waiting_time": 1, "stat_freq": 1,
         "save_freq": 300000, "save_flags": [1, 0, 0, 0], "randomize_v": 1, "randomize_eta": 1,
         "periodic_length": 5.8, "h00": 3.4, "h01": 2.0, "h10": 7.2, "h11": 2.0, "initial_state": 0.0,
         "position0": -1.42, "position1": -0.56, "xc": xc, "randomize_state": 1, "same_initial_barrier": 0,
         "q": 0.2, "tau": 1e-5, "eta0": 0.0, "save_flags2": [0, 0, 0], "rank": 0, "rng": [(0, 0)] * 3}
    R.encode({"header": h, "lanes": lanes, "barrier_state": tab[0][0], "barrier_step": tab[0][1],
              "barrier_min_id": -150, "barrier_state1": tab[1][0], "barrier_step1": tab[1][1],
              "barrier_min_id1": -150, "r_prev": rng.normal(size=W)}, str(tmp_path / "dref_00000_00000.restart"),
             model=1)
    inp = os.path.join(GOLDEN, "cuidb_double_2.0.inp")
    go("import", inp, "dref", "0", str(W), str(tmp_path / "dimg"), "1", "65536")
    with open(tmp_path / "dtab", "wb") as f:
        for st, sp in tab:
            f.write(np.array([-150, n], np.int32).tobytes())
            f.write(np.ascontiguousarray(st.T).tobytes())
            f.write(np.ascontiguousarray(sp.T).tobytes())
    go("export", inp, "1", str(tmp_path / "dimg"), str(tmp_path / "dtab"), "dout", "0", str(W))
    go("import", inp, "dout", "0", str(W), str(tmp_path / "dimg2"), "1", "65536")
    assert (tmp_path / "dimg2").read_bytes() == (tmp_path / "dimg").read_bytes()
    if os.path.isdir(QUICK_JOB):
        go("import", os.path.join(GOLDEN, "quick_test.inp"), os.path.join(QUICK_JOB, "m1.0_t1.00e-05_restart_0"),
           "2", "4", str(tmp_path / "img"))


def _double_lane(rng, W_i, nsteps, nf, cg_sizes, xc):
    """One lane of a double-barrier restart record in the conventions the
    writer follows (previous values equal current ones, flags clear, eta 0)."""
    import numpy as np
    x = float(rng.uniform(-9.0, 9.0))
    periodic = float(np.rint(x * (1.0 / 5.8)))
    well = float(np.floor(x * (1.0 / 5.8)))
    xp = x - periodic * 5.8
    local = 0.0 if xp < xc else 1.0
    state = float(rng.integers(0, 2))
    c0 = int(rng.integers(0, nsteps))
    nfpt = int(rng.integers(0, 40))
    L = {"x": x, "v": 0.0, "f": float(rng.normal()), "e": float(rng.uniform(0, 7)), "eta": 0.0,
         "barrier": float(rng.uniform(0, 7)), "xp": xp,
         "traj_x": rng.normal(size=nf), "current_state": state, "previous_state": state,
         "slope_pos0": 0.0, "slope_pos1": 0.0, "slope_neg0": 0.0, "slope_neg1": 0.0,
         "intercept_pos0": 0.0, "intercept_pos1": 0.0, "intercept_neg0": 0.0, "intercept_neg1": 0.0,
         "current_height0": 0.0, "current_height1": 0.0,
         "state_count0": float(c0), "state_count1": float(nsteps + 1 - c0),
         "well": well, "periodic": periodic, "well_prev": well, "periodic_prev": periodic,
         "entering_well": 0, "entering_period": 0,
         "local": local, "local_prev": local, "entering_local": 0, "local_stored": local,
         "local_prev_stored": local, "entering_local_stored": 0,
         "v0_instance": float(rng.normal()), "eta0_instance": 0.0, "state0_instance": float(rng.integers(0, 2)),
         "well_stored": well, "periodic_stored": periodic, "well_prev_stored": well, "periodic_prev_stored": periodic,
         "entering_well_stored": 0, "entering_period_stored": 0,
         "tick_prev": float(rng.integers(0, nsteps)), "barrier_passed": int(rng.integers(0, 2)),
         "barrier_ticks": np.sort(rng.integers(0, nsteps, int(rng.integers(0, 50)))).astype(np.float64),
         "fpt": rng.exponential(3.0, nfpt), "instant_barrier": rng.integers(0, 2, nfpt).astype(np.float64),
         "initial_barrier": rng.integers(0, 2, nfpt).astype(np.float64),
         "fpt_last_tick": float(rng.uniform(0, 3e5)), "fpt_to_update": int(rng.integers(0, 2)),
         "fpt_init_state": (float(rng.integers(0, 2)), float(rng.integers(0, 2))),
         "fpt_init_index": (int(rng.integers(-5, 5)), int(rng.integers(-5, 5))),
         "fpt2": rng.exponential(30.0, int(rng.integers(0, 60))), "fpt2_last_tick": float(rng.uniform(0, 3e5)),
         "fpt2_to_update": int(rng.integers(0, 2)), "fpt2_index": (int(rng.integers(-5, 5)), int(rng.integers(-5, 5))),
         "tpt": rng.exponential(0.5, int(rng.integers(0, 60))), "tpt_last_tick": float(rng.uniform(0, 3e5)),
         "tpt_to_update": int(rng.integers(0, 2)),
         "tpt2": rng.exponential(0.5, int(rng.integers(0, 60))), "tpt2_last_tick": float(rng.uniform(0, 3e5)),
         "tpt2_to_update": int(rng.integers(0, 2)), "tpt2_initial_well": int(well),
         "cg_sizes": np.array(cg_sizes, np.int32), "cg": []}
    for size in cg_sizes:
        L["cg"].append({"list": rng.exponential(40.0 * size, int(rng.integers(0, 30))),
                        "last_tick": float(rng.uniform(0, 3e5)), "to_update": int(rng.integers(0, 2)),
                        "initial_well": int(rng.integers(-3, 3))})
    for k in ("traj_v", "traj_f", "traj_pe", "traj_ke", "traj_eta", "traj_state", "traj_barrier"):
        L[k] = np.zeros(0)
    return L


def test_double_barrier_reference_layout_round_trip(harness, tmp_path):
    """The double-barrier program's own restart layout
    (langevinDynamicsDoubleBarrier.h:414-642 and 1202-1294; two barrier tables,
    the marker 9948, R_prev): a file in that layout (written by the test's
    encoder) imports into a checkpoint image holding every lane's state,
    local-barrier index, running sums and both barrier arrays; the image
    written back by the program's writer decodes to the same record, lane by
    lane, with the potential constants the reference derives
    (Double.h:1845-1899) continuous at every segment end; that file, read
    back, gives the identical image (taken from after its end marker), and a
    damaged table marker is refused."""
    import numpy as np
    import restart_decode as R
    rng = np.random.default_rng(3)
    W, rank, nsteps, nf = 4, 1, 3 * 10**9, 11
    cg_sizes = [2, 3, 4, 5, 10, 100, 1000]
    x1b, w1, x2b, w2, y1, y2 = -1.42, 0.70, -0.56, 0.65, 3.40, 7.20
    x1h, x2l = x1b + w1, x2b - w2
    s1, s2 = -y1 / (x1h - x1b), y2 / (x2b - x2l)          # falling and rising inner slopes (state 0)
    xc_exact = (x1h * s1 - x2l * s2) / (s1 - s2)
    header = {"signature": 2345, "simd_width": W, "x0": -2.9, "v0": 1.0, "gamma": 1.0, "kb": 1.0,
              "temperature": 1.0, "mass": 1.0, "timestep": 1e-4, "nsteps_total": nsteps, "waiting_time": 1,
              "stat_freq": 1, "save_freq": 300000, "save_flags": [1, 0, 0, 0], "randomize_v": 1, "randomize_eta": 1,
              "periodic_length": 5.8, "h00": 3.4, "h01": 2.0, "h10": 7.2, "h11": 2.0, "initial_state": 0.0,
              "position0": x1b, "position1": x2b, "xc": xc_exact, "randomize_state": 1, "same_initial_barrier": 0,
              "q": 0.2, "tau": 1e-5, "eta0": 0.0, "save_flags2": [0, 0, 0], "rank": rank,
              "rng": [(nsteps, 0), (nsteps, 0), (0, 0)]}
    lanes = [_double_lane(rng, i, nsteps, nf, cg_sizes, xc_exact) for i in range(W)]
    tables = []
    for a in range(2):
        n = 20001
        tables.append((rng.integers(0, 2, (n, W)).astype(np.int32), rng.integers(0, nsteps, (n, W)).astype(np.uint64)))
    d = {"header": header, "lanes": lanes, "barrier_state": tables[0][0], "barrier_step": tables[0][1],
         "barrier_min_id": -10000, "barrier_state1": tables[1][0], "barrier_step1": tables[1][1],
         "barrier_min_id1": -10000, "r_prev": rng.normal(size=W)}
    R.encode(d, str(tmp_path / f"ref_{rank:05d}_00000.restart"), model=1)
    assert R.decode(str(tmp_path / f"ref_{rank:05d}_00000.restart"), model=1)["trailer"] is None
    inp = os.path.join(GOLDEN, "cuidb_double_2.0.inp")

    # import: the reference's record -> checkpoint image
    r = subprocess.run([str(harness), "import", inp, "ref", str(rank), str(W), str(tmp_path / "img"), "1", "65536"],
                       cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert int(r.stdout) == nsteps
    image = (tmp_path / "img").read_bytes()
    hd, got, wbase, win, frames, ticks = parse_image(image)
    assert (hd.model, hd.overdamped, hd.steps_done, hd.n_frames, hd.events_complete) == (1, 1, nsteps, nf, 1)
    for i, L in enumerate(lanes):
        G = got[i]
        assert (G.x, G.xp, G.local_barrier, G.state, G.well, G.periodic) == \
            (L["x"], L["xp"], L["local"], L["current_state"], L["well"], L["periodic"])
        assert (G.fpt2_n, G.tpt_n, G.tpt2_n, G.wt_ticks) == (len(L["fpt2"]), len(L["tpt"]), len(L["tpt2"]),
                                                             len(L["barrier_ticks"]))
        assert G.r_prev == d["r_prev"][i]
        np.testing.assert_array_equal(frames[0, :, i], L["traj_x"])
        for a in range(2):
            for k in range(20001):
                slot = -10000 + k - wbase[i]
                if 0 <= slot < hd.barrier_window:
                    e = int(win[i, a, slot])
                    assert e >> 63 == 1 and (e >> 62) & 1 == tables[a][0][k, i]
                    assert e & ((1 << 62) - 1) == tables[a][1][k, i]

    # export: the image in the reference's layout, with the same tables
    with open(tmp_path / "tables", "wb") as f:
        for st, sp in tables:
            f.write(np.array([-10000, st.shape[0]], np.int32).tobytes())
            f.write(np.ascontiguousarray(st.T).tobytes())
            f.write(np.ascontiguousarray(sp.T).tobytes())
    r = subprocess.run([str(harness), "export", inp, "1", str(tmp_path / "img"), str(tmp_path / "tables"), "out",
                        str(rank), str(W)], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    e = R.decode(str(tmp_path / f"out_{rank:05d}_00000.restart"), model=1)
    assert e["trailer"] == image
    assert abs(e["header"]["xc"] - xc_exact) < 1e-12
    assert {k: v for k, v in e["header"].items() if k != "xc"} == {k: v for k, v in header.items() if k != "xc"}
    xc = e["header"]["xc"]
    yc = (x1h - xc) / (x1h - x1b) * y1
    derived = {k for k in lanes[0] if k.startswith(("slope_", "intercept_", "current_height"))}
    for L, M in zip(lanes, e["lanes"]):
        for k, v in L.items():
            if k in derived:
                continue
            if k == "cg":
                for a, b in zip(v, M[k]):
                    assert np.array_equal(a["list"], b["list"])
                    assert (a["last_tick"], a["to_update"], a["initial_well"]) == \
                        (b["last_tick"], b["to_update"], b["initial_well"])
            elif isinstance(v, np.ndarray):
                np.testing.assert_array_equal(M[k], v, err_msg=k)
            else:
                assert tuple(np.atleast_1d(M[k])) == tuple(np.atleast_1d(v)), k
        # E = xp * slope + intercept on each segment (Double.h:1498-1513): zero at the
        # outer feet, the barrier heights at the barrier tops, yc at the apex xc
        s = int(L["current_state"])
        h0, h1 = (3.4, 2.0)[s], (7.2, 2.0)[s]
        assert (M["current_height0"], M["current_height1"]) == (h0, h1)
        for slope, icpt, x, want in (("slope_neg0", "intercept_neg0", x1b - w1, 0.0),
                                     ("slope_neg0", "intercept_neg0", x1b, h0),
                                     ("slope_pos0", "intercept_pos0", x1b, h0),
                                     ("slope_pos0", "intercept_pos0", xc, yc),
                                     ("slope_neg1", "intercept_neg1", xc, yc),
                                     ("slope_neg1", "intercept_neg1", x2b, h1),
                                     ("slope_pos1", "intercept_pos1", x2b, h1),
                                     ("slope_pos1", "intercept_pos1", x2b + w2, 0.0)):
            term = abs(x * M[slope]) + abs(M[icpt])
            assert abs(x * M[slope] + M[icpt] - want) <= 8 * np.finfo(float).eps * term, (slope, x, s)
    for key in ("barrier_state", "barrier_step", "barrier_state1", "barrier_step1", "r_prev"):
        np.testing.assert_array_equal(e[key], d[key], err_msg=key)
    assert (e["barrier_min_id"], e["barrier_min_id1"]) == (-10000, -10000)

    # this program's reader takes the exact image after the end marker
    r = subprocess.run([str(harness), "import", inp, "out", str(rank), str(W), str(tmp_path / "img2"), "1", "65536"],
                       cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "img2").read_bytes() == image

    # a damaged marker after the two tables is refused
    buf = bytearray((tmp_path / f"ref_{rank:05d}_00000.restart").read_bytes())
    pos = len(buf) - 4 - 8 * W - 4
    assert struct.unpack_from("<i", buf, pos)[0] == 9948
    buf[pos:pos + 4] = struct.pack("<i", 9949)
    (tmp_path / f"ref_{rank:05d}_00000.restart").write_bytes(bytes(buf))
    r = subprocess.run([str(harness), "import", inp, "ref", str(rank), str(W), str(tmp_path / "img3"), "1", "65536"],
                       cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2 and "Error: wrong signiture" in r.stderr
