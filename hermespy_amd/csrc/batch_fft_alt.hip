// Heterogeneous-batch drop kernels with the alternative FFT values per thread (a per-class tuning
// choice of hpx_uneven_prepare), compiled from the fused drop kernel's source in a translation
// unit of their own so that they build in parallel with it.  Only the kernels and the launcher
// hpx::batch_launch_multi_alt come from this unit; every entry point lives in batch.hip.
#define HPX_BATCH_FFT_ALT 1
#include "batch.hip"
