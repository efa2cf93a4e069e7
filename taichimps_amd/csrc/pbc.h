// The periodic wrap of one atom's FP64 position (and the velocity shift of
// a remap-v box), shared by the rebuild kernels and the FP32-split redo's
// fused step start. Every expression rounds as written whichever unit
// includes it, so the wrap gives the same bits in both.
#pragma once

#include "tmps_internal.h"

namespace tmps {

// Domain.pbc_wrap + _pbc_vremap_kernel, orthogonal branch (domain.py:502-571)
__device__ __forceinline__ void pbc_atom(int i, double4* xr, double4* vm, const Ctl* ctl) {
#pragma clang fp contract(off)
    const DevBox& B = ctl->box;
    double4 x = xr[i];
    double xs[3] = {x.x, x.y, x.z};
    if (B.vremap) {
        double nimg[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < 3; ++k)
            if (B.per[k] == 1) nimg[k] = floor((xs[k] - B.lo[k]) / B.prd[k]);
        if (nimg[0] != 0.0 || nimg[1] != 0.0 || nimg[2] != 0.0) {
            const double* h = B.h_rate;
            double4 v = vm[i];
            v.x += -(nimg[0] * h[0] + nimg[1] * h[5] + nimg[2] * h[4]);
            v.y += -(nimg[1] * h[1] + nimg[2] * h[3]);
            v.z += -(nimg[2] * h[2]);
            vm[i] = v;
        }
    }
    for (int k = 0; k < 3; ++k) {
        if (B.per[k] == 1) {
            const double sl = (xs[k] - B.lo[k]) / B.prd[k];
            xs[k] = B.lo[k] + (sl - floor(sl)) * B.prd[k];
        }
    }
    x.x = xs[0]; x.y = xs[1]; x.z = xs[2];
    xr[i] = x;
}

}  // namespace tmps
