// Wavefront shade stage, single-view instantiations for four hero wavelengths.
#include "k_wf_shade.inc"

template __global__ void k_wf_shade<false, 0, true>(DScene, DParams, WfQueues, int, WfBatch);
template __global__ void k_wf_shade<false, 1, true>(DScene, DParams, WfQueues, int, WfBatch);
template __global__ void k_wf_shade<false, 2, true>(DScene, DParams, WfQueues, int, WfBatch);
