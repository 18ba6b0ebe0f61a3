// langevinDynamicsSingleBarrier -- drop-in for the reference executable of
// the same name (langevinDynamicsSingleBarrier.cpp), running on one MI355X.
#include "main_common.h"

int main(int argc, char** argv) { return ldpcp_host::program_main(argc, argv, 0); }
