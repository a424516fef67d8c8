#pragma once
#include "drf_internal.hpp"

namespace drf {

// Y = (X0 ? alpha X0 : 0) + beta * ((X1 ? X1 : 0) + gamma * sum_i c_i K_i)
struct LinComb {
    View<double> Y;
    View<const double> X0, X1;
    View<const double> K[6];
    double c[6];
    int nk;
    double alpha, beta, gamma;
};

hipError_t launch_lincomb(int64_t E, int n, const LinComb &lc, hipStream_t s);

} // namespace drf
