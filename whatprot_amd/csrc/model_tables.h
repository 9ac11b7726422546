// Host-side construction of what the kernels read from the sequencing model,
// shared by the classifier and the fitter, and the HIP status macro of the
// C ABI entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "hmm_device.h"
#include "wp_internal.h"

#define WP_HIP(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            wp::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));              \
            return WP_ERR_HIP;                                                             \
        }                                                                                  \
    } while (0)

namespace wp {

// DevModel of a model with T timesteps, C channels, max_num_dyes M and the
// pruning cutoff (DBL_MAX = uncut). p_detach[i] is DecayingRateModel's rate
// at step i (decaying-rate-model.cc:47-50).
void fill_dev_model(const wp_model& m, int T, int C, int M, double cutoff, DevModel* dm);

// BinomialTransition::reserve (binomial-transition.cc:51-67): the Pascal
// recurrence tables [2][C][(M+1)(M+2)/2], dud (q = p_dud) then bleach
// (q = p_bleach); entry (i, j) at i(i+1)/2 + j is C(i,j) (1-q)^j q^(i-j).
void binomial_tables(const wp_model& m, int C, int M, double* out);

// 0 if a gfx950 device is visible, else WP_ERR_NO_DEVICE with the error set.
int require_device();

}  // namespace wp
