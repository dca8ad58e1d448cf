// CPU stand-in for <hip/hip_runtime.h> used only by the kernel simulator
// (tests/cpu_sim).  It lets the unmodified device source of
// hpdex_kernels.hip run on the host, one OS thread per lane, for debugging
// and CPU tests; it is never part of the shipped library.
#pragma once
#include "../sim_runtime.h"
