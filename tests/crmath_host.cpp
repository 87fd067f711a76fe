// Host build of pyfds_amd/csrc/fds_crmath.h for the CPU tests (the same source the device
// kernels use): pow_cr over arrays, loaded with ctypes.
#include "../pyfds_amd/csrc/fds_crmath.h"

extern "C" void crm_pow(const double *x, const double *y, double *out, long n) {
  for (long i = 0; i < n; ++i) out[i] = fds::crm::pow_cr(x[i], y[i]);
}

extern "C" void crm_exp(const double *z, double *out, long n) {
  for (long i = 0; i < n; ++i) out[i] = fds::crm::exp_cr(z[i]);
}
