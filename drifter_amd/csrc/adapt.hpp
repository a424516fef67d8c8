// Error estimator, element marking and modal coefficients: declarations
// shared by adapt.hip (kernels) and adapt_api.cpp (host + C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "drf_internal.hpp"

namespace drf {

// Upper bound on the first pass's workgroups = length of the partial-max
// buffer an estimator owns.
constexpr int kEstimatorPartials = 2048;

// RefinementParams (include/amr/refinement.hpp:63-71), by value.
struct EstimatorParams {
    double refine_threshold, coarsen_threshold, smoothness_factor;
    int max_level, min_level;
};

// Inverse modal Vandermonde, row-major with row stride kMaxN, by value.
struct ModalOps {
    double Vinv[kMaxN * kMaxN];
};

// V[i][p] = sqrt((2p+1)/2) P_p(x_i) on the n LGL nodes and its inverse
// (adapt_api.cpp); returns 0 on success.
int modal_matrices(int n, double *V, double *Vinv);

hipError_t launch_estimate(int n, int64_t E, const double *field, int64_t es, const double *hchar,
                           const int32_t *levels, const EstimatorParams &p, double *partial,
                           double *max_out, double *value, uint8_t *mask, uint8_t *refine, uint8_t *coarsen,
                           uint8_t *action, uint8_t *mask_out, hipStream_t s);
hipError_t launch_modal(int n, int64_t E, const double *u, int64_t es, const ModalOps &ops,
                        double *coef, uint16_t *bf16, int64_t out_stride, int out_offset,
                        double scale, const double *elem_scale, hipStream_t s);

} // namespace drf
