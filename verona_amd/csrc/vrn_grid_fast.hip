// vrn_grid_fast.hip -- the fast FP64 form of the fused sweeps: the kernels of
// vrn_sweeps.hpp compiled with a*b+c contracted to FMA and reciprocal-based
// division where the divisor is bounded away from zero (see fdiv in
// vrn_physics.hpp).  Same algorithm and operation order otherwise; results
// agree with the exact form to rounding (checked against the oracle at the
// FP64 tolerance rather than bit for bit).
#pragma clang fp contract(fast)
#define VRN_NS vrn_fast
#define VRN_FAST 1
#include "vrn_sweeps.hpp"

void vrn_fast_dispatch_stage5(hipStream_t s, const vrn::SweepArgs &a, const double dxs[3], int rec,
                              int riem, void (*timer)(void *, int, int), void *tctx)
{
    vrn_fast::dispatch_stage<5>(s, a, dxs, rec, riem, false, timer, tctx);
}

void vrn_fast_dispatch_stage4(hipStream_t s, const vrn::SweepArgs &a, const double dxs[3], int rec,
                              int riem, void (*timer)(void *, int, int), void *tctx)
{
    vrn_fast::dispatch_stage<4>(s, a, dxs, rec, riem, false, timer, tctx);
}
