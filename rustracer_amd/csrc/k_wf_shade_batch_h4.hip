// Wavefront shade stage, batch instantiations.
#include "k_wf_shade.inc"

template __global__ void k_wf_shade<true, 0, true>(DScene, DParams, WfQueues, int, WfBatch);
template __global__ void k_wf_shade<true, 1, true>(DScene, DParams, WfQueues, int, WfBatch);
template __global__ void k_wf_shade<true, 2, true>(DScene, DParams, WfQueues, int, WfBatch);
