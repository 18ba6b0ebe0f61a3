// restart_file.h -- the runner's restart files, <OUTPUTRESTART>_%05d_%05d.restart
// per (rank, trial), as LangevinEquationSIMDMultiple::WriteRestartFile /
// ReadRestartFile write and check them (langevinDynamicsSingleBarrier.h:1048-1206;
// langevinDynamicsDoubleBarrier.h:1202-1368).
//
// The file keeps the reference's name, its settings header in the same order
// and encoding (int 4 B, size_t / long long 8 B, double 8 B, little-endian) and
// its read-side consistency checks with the same messages.  Two bodies follow
// the header: the reference's own (signature 2345, write_reference_restart_file),
// with the library's checkpoint image after the end marker, or the image alone
// (signature kRestartSignature, write_restart_file), which the reference program
// refuses at its first check.  This program reads both (read_restart_file).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "runner.h"

namespace ldpcp_host {

constexpr int kRestartSignature = 0x4C445043;  // the reference's files carry 2345
constexpr int kRestartEnd = 9990;

std::string restart_file_name(const std::string& base, int rank, int trial);

// WriteRestartFile: settings header, the three noise positions (here the
// Philox step counter), the checkpoint image and the end marker.
bool write_restart_file(const RunnerConfig& c, int simd_width, int rank, int trial, uint64_t nsteps_total,
                        const std::vector<char>& image);

// What a checkpoint image built from a reference file must say about the
// handle that resumes it (ldpcp_run_resume checks these).
struct ImageSpec {
    uint64_t seed = 0;
    uint32_t traj_offset = 0;
    int32_t barrier_window = 0;
    int32_t n_fields = 1;
    uint64_t tick_capacity = 0;
    uint64_t event_capacity = 0;
    uint32_t n_lists = 0;
};

// Whether the image can be written in the reference program's own layout:
// event lists complete (every event since step 0 kept), all waiting-time ticks
// kept.
bool reference_layout_possible(const RunnerConfig& c, const std::vector<char>& image);

// Widens [*lo, *hi] to every barrier id of `array` (0, or 1 for the double
// barrier's second barrier) a trajectory of the image holds in its window.
void barrier_span(const std::vector<char>& image, int array, int* lo, int* hi);

// The ids the reference creates in every barrier array at step 0, which its
// restart files therefore always list: -100 .. 100 for the single barrier
// (langevinDynamicsSingleBarrier.h:1898-1903), -10000 .. 10000 in both arrays
// of the double barrier (langevinDynamicsDoubleBarrier.h:2250-2255).
int step0_barrier_reach(const RunnerConfig& c);

// One barrier array of the reference's restart file (FiniteInifiniteIntegerTickArray1D,
// DataLocator.h:163-169): ids lo .. lo + n - 1, states and steps as
// ldpcp_run_barrier_table returns them ([trajectory][id]).
struct BarrierTable {
    int lo = 0, n = 0;
    std::vector<int32_t> states;
    std::vector<uint64_t> steps;
};

// WriteRestartFile in the reference program's own layout (signature 2345):
// TrajectoryData::WriteRestartFile (langevinDynamicsSingleBarrier.h:385-576;
// langevinDynamicsDoubleBarrier.h:414-642 with its two barrier tables and the
// marker 9948), the overdamped R_prev (Single.h:2776-2785, Double.h:3391-3400),
// readable by the reference program, followed after its end marker by this
// program's checkpoint image, which read_restart_file prefers (exact
// continuation).  tables holds one table per barrier array of the model.
bool write_reference_restart_file(const RunnerConfig& c, int simd_width, int rank, int trial, uint64_t nsteps_total,
                                  const std::vector<char>& image, const std::vector<BarrierTable>& tables);

// ReadRestartFile: false (after the reference's message on stderr) when the
// file is missing or any setting differs from c.  On success image holds the
// checkpoint image and nsteps_total the header's step count.  Besides this
// program's files it reads the reference program's own files of either model
// (signature 2345; body TrajectoryData::WriteRestartFile,
// langevinDynamicsSingleBarrier.h:385-576 / langevinDynamicsDoubleBarrier.h:414-642,
// then the barrier table(s), DataLocator.h:163-169, and the overdamped R_prev):
// their per-lane state, event lists (as running sums and counts), frames and
// barrier tables become a checkpoint image for `spec`, so a reference run
// continues here (with this program's Philox noise from the file's step count on).
bool read_restart_file(const RunnerConfig& c, int simd_width, int rank, int trial, const ImageSpec& spec,
                       std::vector<char>& image, uint64_t* nsteps_total);

constexpr int kReferenceSignature = 2345;

}  // namespace ldpcp_host
