// One fused pass of the render loop per pixel.
#include "rt_kern.h"

// One pass of the reference's render loop (src/main.rs:463-506) fused per
// pixel: render_tiles' sample, accumulate_sample, and (adaptive, once the
// frame's pass count sample_idx + 1 reaches min_samples) mark_converged.  ctl[0] = stopped (set between
// passes by k_progressive_step), ctl[1] = converged pixels, ctl[2] = passes run.
// H4: four hero wavelengths per camera ray.
template <bool KNN, bool H4>
__global__ __launch_bounds__(BLOCK) void k_progressive(DScene S, DCamera cam, DParams P, rt_accum_opts o, uint32_t pass,
                                                       float* acc, uint32_t* ns, float* m2, uint8_t* conv, uint32_t* ctl) {
    if (ctl[0]) return;
    RT_STACK(st);
    if (KNN) { st.heap = LdsHeap::column(); st.has_heap = true; }
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    uint32_t npix = P.width * P.height;
    bool newly = false;
    if (i < npix && !conv[i]) {
        uint32_t row = i / P.width, col = i - row * P.width;
        Rng rng = rng_seed(pixel_seed(i, P.sample_idx, P.seed));
        Ray r = pixel_ray(cam, P, row, col, P.sample_idx, rng);
        V3 c = ray_color<H4>(S, r, (int)P.max_depth, P.bg_scale, P.use_photons != 0, rng, st);
        float* a = acc + 3 * (size_t)i;
        float* v = m2 + 3 * (size_t)i;
        accumulate_pixel(a, ns + i, v, c, o);
        if (o.adaptive && P.sample_idx + 1 >= o.min_samples && pixel_converged(a, ns[i], v, o.threshold)) {
            conv[i] = 1;
            newly = true;
        }
    }
    wave_count(newly ? 1u : 0u, ctl + 1);
}

template __global__ void k_progressive<true, false>(DScene, DCamera, DParams, rt_accum_opts, uint32_t, float*, uint32_t*, float*, uint8_t*, uint32_t*);
template __global__ void k_progressive<false, false>(DScene, DCamera, DParams, rt_accum_opts, uint32_t, float*, uint32_t*, float*, uint8_t*, uint32_t*);
template __global__ void k_progressive<true, true>(DScene, DCamera, DParams, rt_accum_opts, uint32_t, float*, uint32_t*, float*, uint8_t*, uint32_t*);
template __global__ void k_progressive<false, true>(DScene, DCamera, DParams, rt_accum_opts, uint32_t, float*, uint32_t*, float*, uint8_t*, uint32_t*);
