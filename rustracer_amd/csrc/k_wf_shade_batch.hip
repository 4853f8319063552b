// Wavefront shade stage, batch instantiations.
#include "k_wf_shade.inc"

template __global__ void k_wf_shade<true, 0>(DScene, DParams, WfQueues, int, WfBatch);
template __global__ void k_wf_shade<true, 1>(DScene, DParams, WfQueues, int, WfBatch);
template __global__ void k_wf_shade<true, 2>(DScene, DParams, WfQueues, int, WfBatch);
