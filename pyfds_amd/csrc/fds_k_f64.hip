// Step kernels, float64 form (reproduces the reference's float64 arithmetic).
#define FDS_T double
#define FDS_KAHAN false
#define FDS_SET_NAME kernels_f64
#define FDS_FORM_NS f64
#include "fds_kernels.inc"
