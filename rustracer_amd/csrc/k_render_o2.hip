// The plain megakernel at two waves per SIMD (at most 128 VGPRs).
#include "k_render.inc"

template __global__ void k_render<2, true, false>(DScene, DCamera, DParams, const uint8_t*, float*, uint32_t, int);
template __global__ void k_render<2, false, false>(DScene, DCamera, DParams, const uint8_t*, float*, uint32_t, int);
template __global__ void k_render<2, true, true>(DScene, DCamera, DParams, const uint8_t*, float*, uint32_t, int);
template __global__ void k_render<2, false, true>(DScene, DCamera, DParams, const uint8_t*, float*, uint32_t, int);
