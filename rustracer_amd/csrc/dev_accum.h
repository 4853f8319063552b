// Per-pixel accumulation of sample passes: the firefly clamp and Welford
// statistics of accumulate_sample (src/main.rs:603-644) and the convergence
// test of mark_converged (src/main.rs:646-677), one pixel per call, in the
// reference's f32 operation order (unfused, component-wise Vec3 division).
#pragma once
#include "dev_math.h"
#include "../../include/rt_api.h"

namespace rt {

// Color::luminance (src/vec3.rs:33)
RT_D float accum_lum(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// One sample into one pixel's state (sum a[3], count *n, Welford m2[3]).
// The caller has already skipped converged pixels.
RT_D void accumulate_pixel(float* a, uint32_t* n, float* m2, V3 sc, const rt_accum_opts& o) {
    if (!isfinite(sc.x) || !isfinite(sc.y) || !isfinite(sc.z)) return;
    uint32_t old_n = *n;
    V3 A = mk(a[0], a[1], a[2]);
    V3 old_mean = old_n > 0 ? A / (float)old_n : mk(0, 0, 0);
    float old_lum = accum_lum(old_mean);
    if (o.firefly_clamp > 0.0f && old_n >= o.firefly_min_samples) {
        float l = accum_lum(sc);
        float th = fmaxf(o.firefly_clamp * old_lum, o.firefly_abs_floor);
        if (l > th) sc = sc * (th / l);
    }
    A = A + sc;
    a[0] = A.x; a[1] = A.y; a[2] = A.z;
    *n = old_n + 1;
    V3 new_mean = A / (float)(old_n + 1);
    m2[0] = m2[0] + (sc.x - old_mean.x) * (sc.x - new_mean.x);
    m2[1] = m2[1] + (sc.y - old_mean.y) * (sc.y - new_mean.y);
    m2[2] = m2[2] + (sc.z - old_mean.z) * (sc.z - new_mean.z);
}

// Whether an unconverged pixel with n samples is converged now.
RT_D bool pixel_converged(const float* a, uint32_t n, const float* m2, float threshold) {
    if (n == 0) return false;
    float nf = (float)n;
    float nm1 = (float)(n - 1 > 1 ? n - 1 : 1);
    for (int c = 0; c < 3; c++) {
        float m = a[c] / nf;
        if (m < 1e-4f) continue;
        float variance = m2[c] / nm1;
        float std_err = sqrtf(variance / nf);
        if (std_err / m >= threshold) return false;
    }
    return true;
}

}  // namespace rt
