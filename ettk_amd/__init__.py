"""MI355X-native sub-volume alignment (the hot path of ETTK's MPI_Classification) behind a C ABI.

The compute path is libettk_align.so (hand-written HIP kernels for gfx950, C++ host).  This package holds
the ctypes binding (`ettk_amd.lib`), the C-ABI structs (`ettk_amd.abi`) and the synthetic workload generator
(`ettk_amd.synth`).
"""
