// main_common.h -- option table and value extraction shared by the two
// drop-in programs (langevinDynamicsSingleBarrier.cpp:65-188,
// langevinDynamicsDoubleBarrier.cpp:80-85 and :148-157).
#pragma once
#include <stdio.h>
#include <stdlib.h>

#include "input_parser.h"
#include "runner.h"
#include "text_util.h"

namespace ldpcp_host {

inline void add_options(InputParser& ip, int model)
{
    using R = NumberRequirement;
    using T = OptionType;
    using I = InputType;
    ip.AddOption("OUTPUTBASE", "Output base", model == 0 ? "langevinDynamicsSingleBarrier" : "langevinDynamicsDoubleBarrier",
                 R::Equal, 1, T::MustAndOnce, I::Output);
    ip.AddOption("OVERDAMPED", "Whether use overdamped equation", "false", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("X0", "Initial position (x0)", "0.0", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("V0", "Initial velocity (v0)", "1.0", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("GAMMA", "Friction coefficient (Gamma)", "1.0", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("KB", "Boltzman factor", "1.0", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("TEMPERATURE", "Temperature", "1.0", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("MASS", "Mass", "1.0", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("RANDOMIZEV", "Whether randomize initial velocity", "true", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("PERIODICLENGTH", "Periodic length", "10.0", R::Equal, 1, T::MustAndOnce, I::Parameter);
    if (model == 0) {
        ip.AddOption("BARRIERHEIGHTS", "Barrier heights (two)", "2.0 6.0", R::Equal, 2, T::MustAndOnce, I::Parameter);
        ip.AddOption("BARRIERHALFWIDTH", "Barrier half width", "1.0", R::Equal, 1, T::MustAndOnce, I::Parameter);
    } else {
        ip.AddOption("BARRIERHEIGHTS0", "Barrier heights 0 (two)", "2.0 6.0", R::Equal, 2, T::MustAndOnce, I::Parameter);
        ip.AddOption("BARRIERHEIGHTS1", "Barrier heights 1 (two)", "2.0 6.0", R::Equal, 2, T::MustAndOnce, I::Parameter);
        ip.AddOption("BARRIERHALFWIDTH0", "Barrier half width 0", "1.0", R::Equal, 1, T::MustAndOnce, I::Parameter);
        ip.AddOption("BARRIERHALFWIDTH1", "Barrier half width 1", "1.0", R::Equal, 1, T::MustAndOnce, I::Parameter);
        ip.AddOption("BARRIERPOSITION0", "Barrier position 0", "1.0", R::Equal, 1, T::MustAndOnce, I::Parameter);
        ip.AddOption("BARRIERPOSITION1", "Barrier position 1", "2.0", R::Equal, 1, T::MustAndOnce, I::Parameter);
    }
    ip.AddOption("TIMESTEP", "Timestep (dt)", "0.1", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("NUMBEROFSTEPS", "Number of steps", "10000", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("NUMBEROFPREVIOUSSTEPS", "Number of previous steps", "0", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("NUMBEROFRUNS", "Number of runs", "1", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("FREQSAVE", "Save frequency", "1", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("FREQSTAT", "Statistics frequency", "1", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("DISTRIBUTIONV0", "Whether calculate V0 distribution", "false", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("DISTRIBUTIONS0", "Whether calculate state0 distribution", "false", R::Equal, 1, T::MustAndOnce,
                 I::Parameter);
    ip.AddOption("TAU", "Tau of bichotomic noise eta", "0.1", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("INITIALSTATE", "Initial barrier state (0 or 1)", "0", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("RANDOMIZESTATES", "Whether randomize initial states", "true", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("UNIFORMINITALSTATES", "Uniform initial states", "false", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("EQ", "Whether equal prob, constant forward K", "true 1.0", R::Equal, 2, T::MustAndOnce, I::Parameter);
    ip.AddOption("WAITINGTIME", "Whether calculate waiting time", "false", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("FPT", "Whether calculate first passage time", "false", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("TPT", "Whether calculate transition path time", "false", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("FPTCG", "Whether calculate coarse grained first passage time", "false 10", R::Equal, 2, T::Multiple,
                 I::Parameter);
    ip.AddOption("READRESTART", "Whether read restart", "false", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("WRITERESTART", "Whether write restart", "false", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("INPUTRESTART", "Input restart file base", "inputRestart", R::Equal, 1, T::MustAndOnce, I::Input);
    ip.AddOption("OUTPUTRESTART", "Output restart file base", "outputRestart", R::Equal, 1, T::MustAndOnce, I::Output);
    ip.AddOption("WRITEX", "Whether write x trajectory", "false", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("NUMBEROFLOGSAMPLES", "Number of samples point per order", "20", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("RANDOMSEED", "Random number seed", "-1", R::Equal, 1, T::MustAndOnce, I::Parameter);
    ip.AddOption("PRINTTIMER", "Whether print timer", "false", R::Equal, 1, T::MustAndOnce, I::Parameter);
}

inline RunnerConfig read_config(const InputParser& ip, int model)
{
    RunnerConfig c;
    auto v = [&](const char* k, int i = 0) { return ip.GetOptionValue(k)[i]; };
    c.model = model;
    c.output_base = v("OUTPUTBASE");
    c.x0 = to_double(v("X0"));
    c.v0 = to_double(v("V0"));
    c.gamma = to_double(v("GAMMA"));
    c.kb = to_double(v("KB"));
    c.temperature = to_double(v("TEMPERATURE"));
    c.mass = to_double(v("MASS"));
    c.timestep = to_double(v("TIMESTEP"));
    c.nsteps = to_ulong(v("NUMBEROFSTEPS"));
    c.save_freq = to_ulong(v("FREQSAVE"));
    c.stat_freq = to_ulong(v("FREQSTAT"));
    c.randomize_v = to_boolean(v("RANDOMIZEV"));
    c.randomize_state = to_boolean(v("RANDOMIZESTATES"));
    c.same_initial_barrier = to_boolean(v("UNIFORMINITALSTATES"));
    c.periodic_length = to_double(v("PERIODICLENGTH"));
    if (model == 0) {
        c.height[0][0] = c.height[1][0] = to_double(v("BARRIERHEIGHTS", 0));
        c.height[0][1] = c.height[1][1] = to_double(v("BARRIERHEIGHTS", 1));
        c.half_width[0] = c.half_width[1] = to_double(v("BARRIERHALFWIDTH"));
    } else {
        c.height[0][0] = to_double(v("BARRIERHEIGHTS0", 0));
        c.height[0][1] = to_double(v("BARRIERHEIGHTS0", 1));
        c.height[1][0] = to_double(v("BARRIERHEIGHTS1", 0));
        c.height[1][1] = to_double(v("BARRIERHEIGHTS1", 1));
        c.half_width[0] = to_double(v("BARRIERHALFWIDTH0"));
        c.half_width[1] = to_double(v("BARRIERHALFWIDTH1"));
        c.position[0] = to_double(v("BARRIERPOSITION0"));
        c.position[1] = to_double(v("BARRIERPOSITION1"));
    }
    // INITIALSTATE is read as a double and passed on as an int (Single.cpp:143, 3134)
    c.initial_state = (double)(int)to_double(v("INITIALSTATE"));
    c.tau = to_double(v("TAU"));
    c.waiting = to_boolean(v("WAITINGTIME"));
    c.fpt = to_boolean(v("FPT"));
    c.tpt = to_boolean(v("TPT"));
    if (ip.GetOptionSpecified("FPTCG")) {
        for (size_t j = 0; j < ip.GetNumberOfOptionValue("FPTCG"); ++j) {
            const auto w = ip.GetOptionValue("FPTCG", (int)j);
            c.fptcg = to_boolean(w[0]);  // the last entry's flag decides (Single.cpp:156-162)
            c.cg_size.push_back(to_int(w[1]));
        }
    }
    c.read_restart = to_boolean(v("READRESTART"));
    c.write_restart = to_boolean(v("WRITERESTART"));
    c.input_restart = v("INPUTRESTART");
    c.output_restart = v("OUTPUTRESTART");
    c.ntrial = to_int(v("NUMBEROFRUNS"));
    c.nsteps_previous = to_ulong(v("NUMBEROFPREVIOUSSTEPS"));
    c.write_x = to_boolean(v("WRITEX"));
    c.sample_per_order = (int)to_ulong(v("NUMBEROFLOGSAMPLES"));
    c.seed = to_int(v("RANDOMSEED"));
    c.timer = to_boolean(v("PRINTTIMER"));
    c.v0_distribution = to_boolean(v("DISTRIBUTIONV0"));
    c.state0_distribution = to_boolean(v("DISTRIBUTIONS0"));
    c.overdamped = to_boolean(v("OVERDAMPED"));
    c.equal_probability = to_boolean(v("EQ", 0));
    c.eq_constant = to_double(v("EQ", 1));
    c.ranks = ip.ranks;
    c.simd_width = ip.simd_width;
    c.steps_per_launch = ip.steps_per_launch;
    return c;
}

// Under an MPI launcher (mpirun -n R ...), rank 0 runs all R emulated ranks on
// the GPU (their seeds, layout and reductions are the reference's); the other
// launched processes have nothing to do and exit quietly.
inline int launcher_rank()
{
    for (const char* name : {"OMPI_COMM_WORLD_RANK", "PMI_RANK", "PMIX_RANK", "LDPCP_RANK"}) {
        const char* v = getenv(name);
        if (v && *v) return atoi(v);
    }
    return 0;
}

inline int program_main(int argc, char** argv, int model)
{
    if (launcher_rank() > 0) return 0;
    InputParser ip;
    add_options(ip, model);
    if (!ip.Parse(argc, argv, true)) return 0;
    return run(read_config(ip, model));
}

}  // namespace ldpcp_host
