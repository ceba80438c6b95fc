// surrogate.h -- collision-probability MLP (7 -> 256 -> 256 -> 256 -> 1, SiLU,
// sigmoid output), BF16 weights and activations on MFMA with FP32
// accumulation, trained on labels of the exact checker.  It only ranks
// configurations; it never decides one.
#pragma once

#include <hip/hip_runtime.h>

namespace gprm {

struct Surrogate;

// Collision probabilities of n configurations (SoA, component i of item k at
// q[k + offset*i], raw joint angles) into prob[n]; 0 on success.
int surrogate_predict(Surrogate* s, const float* q, int n, long long offset, float* prob, hipStream_t st);

}  // namespace gprm
