"""MI355X-native engine for the p=619 doubly-focused pseudocube sieve.

Drop-in for the batch entry points of AlexanderEastwood/psuedocubes:
`gpu` (Runtime, GPUFocus, Matcher), `batch_experiment` (prepare,
MaskedMatcher, compare), `optimized` (compute_lease, Accumulator) and
`run_production` (compute_batch).  The device work is done by the HIP
kernels of libpcube.so, reached through the C ABI in include/pcube.h.
"""
