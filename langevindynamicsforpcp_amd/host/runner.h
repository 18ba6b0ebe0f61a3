// runner.h -- the drop-in runner: LANGEVINEQUATION::LangevinEquationRunnerMPI
// (langevinDynamicsSingleBarrier.h:2890-4040; langevinDynamicsDoubleBarrier.h:3509-4718)
// on top of the C ABI in include/ldpcp.h.
//
// The reference splits trajectories over MPI ranks of SIMDWIDTH lanes each.
// Here one process drives the GPU; the ranks are emulated ("virtual ranks"):
// rank r keeps its own seed (the r-th distinct glibc rand() after
// srand(RANDOMSEED), mpiKernel.cpp:124-138) as its Philox key, and every
// output keeps the reference's rank-major order.
#pragma once
#include <string>
#include <vector>

namespace ldpcp_host {

struct RunnerConfig {
    int model = 0;  // 0 single, 1 double barrier
    std::string output_base = "langevinEquation";
    double x0 = 0.0, v0 = 0.0, gamma = 1.0, kb = 1.0, temperature = 1.0, mass = 1.0;
    double timestep = 0.1;
    unsigned long nsteps = 10000, stat_freq = 1, save_freq = 1;
    bool randomize_v = false, randomize_state = true, same_initial_barrier = false;
    double periodic_length = 5.0;
    double height[2][2] = {{2.0, 6.0}, {2.0, 6.0}};
    double half_width[2] = {1.0, 1.0};
    double position[2] = {0.0, 0.0};
    double initial_state = 0.0;
    double q = 0.2, tau = 0.1, eta0 = 0.0;
    bool waiting = false, fpt = false, tpt = false, fptcg = false;
    std::vector<int> cg_size;
    bool read_restart = false, write_restart = false;
    std::string input_restart = "inputRestart", output_restart = "outputRestart";
    int ntrial = 1;
    unsigned long nsteps_previous = 0;
    bool write_x = false;
    int sample_per_order = 20;
    int seed = -1;
    bool timer = false;
    bool v0_distribution = false, state0_distribution = false;
    bool overdamped = false;
    bool equal_probability = true;
    double eq_constant = 1.0;
    // emulation of the MPI layout
    int ranks = 1;
    int simd_width = 8;
    unsigned long steps_per_launch = 1ul << 20;
};

// Returns the process exit status (the reference exits 0 on its error paths).
int run(const RunnerConfig& cfg);

// Seeds of the emulated ranks: srand(seed, or time(0) when seed == -1), then
// distinct rand() values (RANDOMNUMBER.cpp:5-34).
std::vector<int> rank_seeds(int master_seed, int ranks, int* used_master);

// MOLUTILITY::GenerateLog10Scale (MOLUTILITY.cpp:59-73)
std::vector<unsigned long> log10_scale(double timestep, unsigned long n_frames, unsigned long minimal,
                                       unsigned long per_order, bool include_zero);

}  // namespace ldpcp_host
