// vrn_tubes_fast.hip -- the fast FP64 form of the batched 1D tubes: the
// kernels of vrn_tubes_kern.hpp compiled with a*b+c contracted to FMA and
// reciprocal-based division (fdiv / recip in vrn_physics.hpp); the 3x3 Newton
// con2prim takes one reciprocal of gamma and of the Jacobian determinant and
// compares the squared step norm with tol^2.  Same algorithm and operation
// order otherwise; results agree with the exact form to rounding on
// well-conditioned states.
#pragma clang fp contract(fast)
#define VRN_NS vrn_fast
#define VRN_FAST 1
#include "vrn_tubes_kern.hpp"

void vrn_fast_launch_tubes(hipStream_t s, const vrn::TubeArgs &a, int rec, int riem, int block, int grid)
{
    vrn_fast::launch_tubes_kernels(s, a, rec, riem, block, grid);
}
