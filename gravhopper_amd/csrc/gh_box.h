// Root box of a point set from per-workgroup extents, shared by the box kernel of the
// tree build (gh_tree.hip) and the key kernel of the small spatial sort (gh_frames.hip),
// so that both form the same box bit for bit.
#pragma once

#include "gh_internal.h"

namespace gh {

// Minimum and maximum of each coordinate over the workgroup's threads: smin/smax hold
// every thread's values on entry and the result in entry 0 on exit.
__device__ __forceinline__ void block_extent(double (*smin)[kBlock], double (*smax)[kBlock])
{
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 3; k++) {
                if (smin[k][threadIdx.x + w] < smin[k][threadIdx.x])
                    smin[k][threadIdx.x] = smin[k][threadIdx.x + w];
                if (smax[k][threadIdx.x + w] > smax[k][threadIdx.x])
                    smax[k][threadIdx.x] = smax[k][threadIdx.x + w];
            }
        __syncthreads();
    }
}

// The reference's root box (gravhopper/_jbgrav.c:764-772) of the nparts extents
// part[6 b ..] = (min x, y, z, max x, y, z), returned to every thread of the workgroup:
// box[0..2] = centre, box[3] = size.  Minima and maxima are exact in any order.
__device__ __forceinline__ void box_from_parts(const double* __restrict__ part, int nparts,
                                               double eps, double (*smin)[kBlock],
                                               double (*smax)[kBlock], double box[4])
{
    double mn[3], mx[3];
    for (int k = 0; k < 3; k++) {
        mn[k] = part[k];
        mx[k] = part[3 + k];
    }
    for (int b = threadIdx.x; b < nparts; b += kBlock)
        for (int k = 0; k < 3; k++) {
            if (part[6 * b + k] < mn[k]) mn[k] = part[6 * b + k];
            if (part[6 * b + 3 + k] > mx[k]) mx[k] = part[6 * b + 3 + k];
        }
    for (int k = 0; k < 3; k++) {
        smin[k][threadIdx.x] = mn[k];
        smax[k][threadIdx.x] = mx[k];
    }
    block_extent(smin, smax);
    for (int k = 0; k < 3; k++) {
        mn[k] = smin[k][0];
        mx[k] = smax[k][0];
    }
    double bs = mx[0] - mn[0] + eps;
    for (int k = 1; k < 3; k++)
        if ((mx[k] - mn[k]) > bs) bs = mx[k] - mn[k] + eps;
    for (int k = 0; k < 3; k++) box[k] = 0.5 * (mn[k] + mx[k]);
    box[3] = bs;
    __syncthreads();  // smin/smax may be reused
}

}  // namespace gh
