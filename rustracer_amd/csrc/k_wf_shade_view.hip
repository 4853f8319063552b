// Wavefront shade stage, single-view instantiations.
#include "k_wf_shade.inc"

template __global__ void k_wf_shade<false, 0>(DScene, DParams, WfQueues, int, WfBatch);
template __global__ void k_wf_shade<false, 1>(DScene, DParams, WfQueues, int, WfBatch);
template __global__ void k_wf_shade<false, 2>(DScene, DParams, WfQueues, int, WfBatch);
