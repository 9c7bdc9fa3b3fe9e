// Host-side pieces of the Jacobi sweeps shared by the plan entry points and
// the no-allocation entry point: the kernel argument block, the launch
// configuration of one parameter set, and the enqueue of one run.
#pragma once

#include "cfm_common.hpp"

namespace cfm {

struct JacobiArgs {
    void *buf0;
    void *buf1;
    int nx, ny, nz;
    int niter;        // sweeps in this launch
    int it0;          // index of the first sweep (parity and norm slots)
    int bc;
    int norm_every;
    int chunks;       // chunks along the streamed axis
    int chunk_len;
    int ntx, nty;     // tiles along x and y (nty = 1 in 2D)
    int no_compute;
    double *norm_part;  // [slot][gridDim.x]
    GridSync *sync;
    unsigned long long *timers;  // phase timers: [workgroup][kTimerSlots] ticks, or null
};

using KernelFn = void (*)(JacobiArgs);

// Launch configuration of one parameter set: kernel, workgroup, grid and the
// work split.  Computed without allocating anything (device queries only).
struct JacobiConfig {
    DeviceProps dev{};
    size_t elems = 0;
    size_t elem_size = 0;
    int norm_slots = 0;
    KernelFn kernel = nullptr;
    bool t2 = false;  // more than one sweep per pass (the two-/three-/four-sweep kernels)
    int sweeps_per_pass = 1;
    int grid = 0;
    int chunks = 0, chunk_len = 0, ntx = 0, nty = 0, rows = 0, block = 256;
    unsigned long long spin_ticks = 0;
};

// Validates p and fills c for the current device.
cfm_status jacobi_configure(const cfm_jacobi_params *p, JacobiConfig *c);
// Enqueues one run of p.niter sweeps on `stream` (barrier reset, then the
// persistent launch or one launch per sweep) over b0 (current field) and b1;
// timers (may be null): per-workgroup phase ticks; *coop as launch_persistent
// reports it (-1 without a persistent launch).
cfm_status jacobi_enqueue(const cfm_jacobi_params &p, const JacobiConfig &c, void *b0, void *b1, GridSync *sync,
                          double *norm_part, unsigned long long *timers, hipStream_t stream, int *coop);
// Passes over HBM of one run (its parity says which array holds the result).
int jacobi_passes(const cfm_jacobi_params &p, const JacobiConfig &c);
void jacobi_fill_timing(cfm_timing *t, const cfm_jacobi_params &p, const JacobiConfig &c, float ms, double wall_ms,
                        int coop);

}  // namespace cfm
