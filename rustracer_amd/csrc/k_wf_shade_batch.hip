// Wavefront shade stage, batch instantiations.
#include "k_wf_shade.inc"

template __global__ void k_wf_shade<true, 0, false>(DScene, DParams, WfQueues, int, WfBatch);
template __global__ void k_wf_shade<true, 1, false>(DScene, DParams, WfQueues, int, WfBatch);
template __global__ void k_wf_shade<true, 2, false>(DScene, DParams, WfQueues, int, WfBatch);
