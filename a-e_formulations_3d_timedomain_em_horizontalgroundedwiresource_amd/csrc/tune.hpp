// Autotuner of the solver's launch parameters (tile shape choice, outer block, look-ahead LDS,
// sweep strategy) by timing candidates on the model's own matrix.
#pragma once

#include <functional>
#include <string>

#include "chol.hpp"
#include "symbolic.hpp"

namespace emtd {

// Coordinate descent over the launch parameters, starting from the solver's current tuning:
// factorization parameters first (outer block 64/128/256, 128x128 tiles from 128/256/512 target
// columns, look-ahead LDS padding on/off, XCD-grouped update tile order on/off), each candidate
// timed by refactoring the same matrix
// (refactor() returns false on a pivot failure, which rejects the candidate); then the sweep
// strategy (split-sweep threshold 0/256/1024/4096 fronts per level, one-wavefront-per-front
// sweeps on/off), each timed over three solves of b into x. A candidate replaces the best one
// only if it is at least 2% faster. Leaves the solver with the best tuning and its factor.
template <typename T>
Tuning autotune(GpuCholesky<T>& C, const std::function<bool()>& refactor, const double* d_b, double* d_x, int nrhs,
                int64_t ld, std::string* report);

// Process-wide cache of tuned parameters, keyed by the factor structure and the precision
// (tuning_key), or by a size class: precision, number of levels and the half-octave of the
// factored DOF count (tuning_class_key), which models of similar size share.
std::string tuning_key(const Symbolic& S, int elem_bytes);
std::string tuning_class_key(const Symbolic& S, int elem_bytes);
bool tuning_cache_get(const std::string& key, Tuning* t);
void tuning_cache_put(const std::string& key, const Tuning& t);

}  // namespace emtd
