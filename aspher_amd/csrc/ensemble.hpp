// Ensembles: many independent certified contact solves of uneven grid sizes
// advanced together on one GPU.
//
// The instances' vectors are concatenated in HBM; every O(N) pass of an
// iteration is ONE launch over a per-instance work queue (chunks of every
// active instance), every reduction is finished for all instances by one
// launch, and the spectral operators run one batched launch per group of
// equal (kind, Ns, L, E*). The device decides convergence, best iterate and
// stall per instance (finalize step STEP_B_EVAL2) and drops finished
// instances from the work, so an iteration has no host round trip: it is
// captured once in a HIP graph and replayed; the host only polls the
// instances' flags every few replays. Instances that need a host decision
// (identification step, stall, budget) finish on the single-solve path,
// warm-started from their device state.
#pragma once

#include "engine.hpp"

#include <string>
#include <vector>

namespace aspher {

struct EnsembleInstance {
    int kind = 0;              // 0 free space ("fft"), 1 periodic
    int Ns = 0;
    double L = 1.0, E_star = 1.0;
    double p_bar = 0.0;
    const double* gap = nullptr;   // device FP64, Ns^2 (caller-owned)
};

struct EnsembleOptions {
    double tol = 1e-5;
    int max_iter = 5000;
    bool use_pr = true;
    std::string precision = "float";   // "float" (FP32 + FP64 reductions) | "double"
    std::string precond = "auto";      // auto | none | fourier | fourier-fft
    int check_every = 8;               // graph replays between host polls
    bool use_graph = true;
    bool autotune = true;
    bool keep_best = true;
    int stall_limit = 200;
    bool light = false;
};

struct EnsembleStats {
    double seconds = 0.0, setup_seconds = 0.0, loop_seconds = 0.0, finish_seconds = 0.0;
    double tune_seconds = 0.0;
    int groups = 0, replays = 0, polls = 0, host_fallbacks = 0, device_iterations = 0;
    bool graph = false;
    std::vector<double> instance_seconds;   // loop time until each instance left the device
};

// Solves every instance; results in input order.
void solve_ensemble(const std::vector<EnsembleInstance>& inst, const EnsembleOptions& opt,
                    std::vector<Result>& out, EnsembleStats& stats);

// Launch configuration of a batched spectral plan chosen by timing the
// candidate tile shapes on the device (process-wide cache).
template <typename T>
SpectralTuning autotune_plan(int n, int rows_in, int cols_in, int rows_out, int cols_out,
                             int batch, Symbol sym, hipStream_t s);

} // namespace aspher
