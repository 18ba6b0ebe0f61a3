"""Host helpers of the drop-in programs, on the CPU.

Output backups follow MOLUTILITY::findNextBackUpFileName (MOLUTILITY.cpp:6-22):
an existing file `dir/name` is renamed to `dir/#name_bk.N#` with the first
free N before a new file of that name is written.  Value conversions follow
the reference's `text` class (text.cpp:215-308).
"""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(REPO, "langevindynamicsforpcp_amd", "host")

HARNESS = r"""
#include <stdio.h>
#include <string.h>
#include "text_util.h"
using namespace ldpcp_host;
int main(int argc, char** argv)
{
    if (!strcmp(argv[1], "backup")) { backup_existing(argv[2]); return 0; }
    if (!strcmp(argv[1], "bool")) { printf("%d\n", to_boolean(argv[2]) ? 1 : 0); return 0; }
    if (!strcmp(argv[1], "split")) {
        for (const auto& w : split(remove_comment(argv[2], "#"))) printf("[%s]", w.c_str());
        printf("\n");
        return 0;
    }
    return 1;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("text")
    (d / "h.cpp").write_text(HARNESS)
    exe = d / "h"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", HOST, str(d / "h.cpp"),
                           os.path.join(HOST, "text_util.cpp"), "-o", str(exe)])
    return exe


def test_backup_names_follow_the_reference(harness, tmp_path):
    sub = tmp_path / "out"
    sub.mkdir()
    target = sub / "run_msd.log"
    for k in range(3):
        target.write_text(f"v{k}")
        subprocess.check_call([str(harness), "backup", str(target)])
        assert not target.exists()
    assert sorted(os.listdir(sub)) == ["#run_msd.log_bk.1#", "#run_msd.log_bk.2#", "#run_msd.log_bk.3#"]
    assert (sub / "#run_msd.log_bk.2#").read_text() == "v1"
    # a gap is filled first, as the reference's search from N = 1 does
    (sub / "#run_msd.log_bk.1#").unlink()
    target.write_text("v3")
    subprocess.check_call([str(harness), "backup", str(target)])
    assert (sub / "#run_msd.log_bk.1#").read_text() == "v3"
    # nothing to back up: nothing happens
    subprocess.check_call([str(harness), "backup", str(sub / "absent.dat")])
    assert len(os.listdir(sub)) == 3


@pytest.mark.parametrize("word,want", [("true", 1), ("TRUE", 1), ("True", 1), ("false", 0), ("0", 0), ("1", 1),
                                       ("2.5", 1), ("yes", 0)])
def test_to_boolean(harness, word, want):
    out = subprocess.check_output([str(harness), "bool", word], text=True)
    assert int(out) == want


def test_split_and_comment(harness):
    out = subprocess.check_output([str(harness), "split", "FPTCG   true\t10  # coarse grain"], text=True)
    assert out.strip() == "[FPTCG][true][10]"
