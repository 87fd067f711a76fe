// Step kernels, float32 form with Kahan-compensated accumulation of the updated fields.
#define FDS_T float
#define FDS_KAHAN true
#define FDS_SET_NAME kernels_k32
#define FDS_FORM_NS k32
#include "fds_kernels.inc"
