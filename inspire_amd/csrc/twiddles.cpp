// FFT twiddle table of crypto/fft.h:71-80, compiled with g++ (not hipcc).
//
// The reference computes twiddles_forward_[i] = (cos(a), sin(a)) and
// twiddles_backward_[i] = (cos(-a), sin(-a)) with a = -2*pi*i/N, and its
// rounded FFT products (crypto/polynomial_fft.cc:183-186) can depend on the
// last bit of these doubles. GCC lowers this expression to glibc's sincos()
// plus sin(), while clang (hipcc) lowers it to separate sin()/cos() calls; the
// two differ by one ulp for a few angles. This file is therefore built by the
// same compiler family as the reference so the table is bit-identical to it.
#include <cmath>
#include <cstddef>

extern "C" void pirg_internal_twiddles(int n, double* fwd, double* bwd) {
  const double pi = std::acos(-1.0);
  const size_t size = static_cast<size_t>(n);
  for (size_t i = 0; i < size / 2; ++i) {
    double angle = -2.0 * pi * i / size;
    fwd[2 * i] = std::cos(angle);
    fwd[2 * i + 1] = std::sin(angle);
    bwd[2 * i] = std::cos(-angle);
    bwd[2 * i + 1] = std::sin(-angle);
  }
}
