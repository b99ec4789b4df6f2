// cih_blur.hpp -- one extended-box pass of the shadow blur over a line of
// samples (canvas_ity.hpp:2460-2481 along rows, 2483-2503 along columns).
//
// The blur kernels run one thread per line with the lines of a wave side by
// side in memory (sample x of line t at base + x * step + t), so every load and
// store of a wave is one contiguous 256-byte access.  The running sum is the
// reference's: the same products added and subtracted in the same order under
// the same guards, so the result is bit-identical to a sequential pass.
#ifndef CIH_BLUR_HPP
#define CIH_BLUR_HPP

#include <stddef.h>

#include "cih_geometry.hpp"

namespace cih {

// Steps of the steady-state middle taken at a time, their loads issued first.
#define CIH_BLUR_BLOCK 8

// out[x * step] for x in [0, n): the window sum over in[x - r - 2 .. x + r + 1]
// with weight w2 on the inner 2r + 1 samples and w1 on the two outer ones,
// updated incrementally from x - 1 as the reference does.  Samples past the
// end of the line count as zero in the first window (the reference reads its
// zero-initialised scratch there); later terms past either end are skipped.
// Each sample is loaded twice, as the leading edge (x + r + 1) and, 2r + 2
// steps later, as the trailing edge (x - r - 1); the values one step old are
// carried in registers.
CIH_HD void box_line(const float *__restrict__ in, float *__restrict__ out, int n, size_t step,
                     int r, float w1, float w2) {
    if (n <= 0)
        return;
    float running = w1 * (r + 1 < n ? in[(size_t)(r + 1) * step] : 0.0f);
    for (int x = 0; x <= r; ++x)
        running += (w1 + w2) * (x < n ? in[(size_t)x * step] : 0.0f);
    out[0] = running;
    // lead = in[x + r] and trail = in[x - r - 2] for the next x, where valid
    float lead = r + 1 < n ? in[(size_t)(r + 1) * step] : 0.0f, trail = 0.0f;
    int x = 1;
    // ramp-up: the window's trailing edge is still before the line
    for (; x < n && x < r + 2; ++x) {
        float a = x >= r + 1 ? in[(size_t)(x - r - 1) * step] : 0.0f;
        float e = x + r + 1 < n ? in[(size_t)(x + r + 1) * step] : 0.0f;
        if (x >= r + 1)
            running -= w2 * a;
        if (x + r < n)
            running += w2 * lead;
        if (x + r + 1 < n)
            running += w1 * e;
        out[(size_t)x * step] = running;
        trail = a;
        lead = e;
    }
    // steady state: all four terms, CIH_BLUR_BLOCK steps per block
    int middle = n - r - 1; // x + r + 1 < n
    for (; x + CIH_BLUR_BLOCK <= middle; x += CIH_BLUR_BLOCK) {
        float a[CIH_BLUR_BLOCK], e[CIH_BLUR_BLOCK];
#pragma unroll
        for (int i = 0; i < CIH_BLUR_BLOCK; ++i) {
            a[i] = in[(size_t)(x + i - r - 1) * step];
            e[i] = in[(size_t)(x + i + r + 1) * step];
        }
#pragma unroll
        for (int i = 0; i < CIH_BLUR_BLOCK; ++i) {
            running -= w2 * a[i];
            running -= w1 * trail;
            running += w2 * lead;
            running += w1 * e[i];
            out[(size_t)(x + i) * step] = running;
            trail = a[i];
            lead = e[i];
        }
    }
    // the rest, leading edge running off the end
    for (; x < n; ++x) {
        float a = in[(size_t)(x - r - 1) * step];
        float e = x + r + 1 < n ? in[(size_t)(x + r + 1) * step] : 0.0f;
        running -= w2 * a;
        running -= w1 * trail;
        if (x + r < n)
            running += w2 * lead;
        if (x + r + 1 < n)
            running += w1 * e;
        out[(size_t)x * step] = running;
        trail = a;
        lead = e;
    }
}

} // namespace cih

#endif
