// fft_lines.h -- radix-2 decimation-in-time transforms of lines held in LDS
// (the transform of the CPU restatement, stage for stage), shared by the 2D
// solves and the 1D electromagnetic solve.
#pragma once
#include "pic_device.h"

namespace pic {

// ---------------------------------------------------------------------------
// Line transforms: `rows` lines of length L = 2^lL in LDS, line r at r*stride,
// element i at padi(i).  Each pass applies up to three radix-2 stages to
// groups held in registers (radix2_group).
// ---------------------------------------------------------------------------
template <int M>
__device__ __forceinline__ void fft_lines_pass(double* re, double* im, const double2* __restrict__ tw, int lL,
                                               int rows, int stride, int s0, bool inverse) {
    constexpr int G = 1 << M;
    const int gl = lL - M;  // log2 groups per line
    const int groups = rows << gl;
    const double2* tb = tw + ((1 << s0) - 1);
    for (int gi = threadIdx.x; gi < groups; gi += blockDim.x) {
        const int r = gi >> gl;
        const int g = gi & ((1 << gl) - 1);
        const int lo = g & ((1 << s0) - 1);
        const int hi = g >> s0;
        const int base = (hi << (s0 + M)) + lo;
        double* lr = re + r * stride;
        double* li = im + r * stride;
        double ar[G], ai[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int id = padi(base + (j << s0));
            ar[j] = lr[id];
            ai[j] = li[id];
        }
        radix2_group<M>(ar, ai, tb, lo, s0, inverse);
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int id = padi(base + (j << s0));
            lr[id] = ar[j];
            li[id] = ai[j];
        }
    }
    __syncthreads();
}

__device__ inline void fft_lines(double* re, double* im, const double2* __restrict__ tw, int lL, int rows, int stride,
                          bool inverse) {
    for (int s0 = 0; s0 < lL; s0 += kPassStages) {
        const int m = lL - s0 < kPassStages ? lL - s0 : kPassStages;
        if (m == 3) fft_lines_pass<3>(re, im, tw, lL, rows, stride, s0, inverse);
        else if (m == 2) fft_lines_pass<2>(re, im, tw, lL, rows, stride, s0, inverse);
        else fft_lines_pass<1>(re, im, tw, lL, rows, stride, s0, inverse);
    }
}

}  // namespace pic
