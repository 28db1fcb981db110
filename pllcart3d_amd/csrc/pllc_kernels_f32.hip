// pllc_kernels_f32.hip -- FP32 instantiation of the substage kernels (mixed
// path), built with contraction and fast single-precision division; the
// reductions inside these kernels still accumulate in FP64.
#include "pllc_kernels_impl.h"
