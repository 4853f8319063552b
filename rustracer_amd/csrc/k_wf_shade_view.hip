// Wavefront shade stage, single-view instantiations.
#include "k_wf_shade.inc"

template __global__ void k_wf_shade<false, 0, false>(DScene, DParams, WfQueues, int, WfBatch);
template __global__ void k_wf_shade<false, 1, false>(DScene, DParams, WfQueues, int, WfBatch);
template __global__ void k_wf_shade<false, 2, false>(DScene, DParams, WfQueues, int, WfBatch);
