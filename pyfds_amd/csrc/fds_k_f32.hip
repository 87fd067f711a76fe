// Step kernels, plain float32 form.
#define FDS_T float
#define FDS_KAHAN false
#define FDS_SET_NAME kernels_f32
#define FDS_FORM_NS f32
#include "fds_kernels.inc"
