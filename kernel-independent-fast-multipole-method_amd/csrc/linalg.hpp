// linalg.hpp -- host FP64 numerics used by operator precompute.
#pragma once

#include <vector>

namespace kifmm {

// Truncated pseudoinverse of the m x n column-major matrix A, keeping the
// singular values sigma_i >= eps * sigma_max.  R becomes n x m column-major.
// Returns the kept rank.
int pinv_truncated(int m, int n, const double* A, double eps, std::vector<double>& R);

// Real-to-complex 3-D DFT of an n^3 array laid out [z][y][x] (x fastest);
// out is [kz][ky][kx] complex (re, im interleaved), kx in [0, n/2].
// Sign -1, unnormalised.
void rdft3(int n, const double* in, double* out);

}  // namespace kifmm
