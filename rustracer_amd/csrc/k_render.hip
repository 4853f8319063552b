// The plain megakernel at the compiler's register budget.
#include "k_render.inc"

template __global__ void k_render<0, true, false>(DScene, DCamera, DParams, const uint8_t*, float*, uint32_t, int);
template __global__ void k_render<0, false, false>(DScene, DCamera, DParams, const uint8_t*, float*, uint32_t, int);
template __global__ void k_render<0, true, true>(DScene, DCamera, DParams, const uint8_t*, float*, uint32_t, int);
template __global__ void k_render<0, false, true>(DScene, DCamera, DParams, const uint8_t*, float*, uint32_t, int);
