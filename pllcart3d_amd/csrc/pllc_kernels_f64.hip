// pllc_kernels_f64.hip -- FP64 instantiation of the substage kernels (plain
// path), built with -ffp-contract=off: no multiply-add is fused, so every
// kernel repeats the reference's FP64 arithmetic operation by operation.
#define PLLC_F64_TU
#include "pllc_kernels_impl.h"
