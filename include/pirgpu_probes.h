/*
 * pirgpu_probes — calibration kernels for measurement (libpirgpu_probes.so).
 * Not part of the answer path or of the PirServer interface: the benchmark
 * and the profiling runs use them to scale HBM byte counters and to place the
 * answer kernels against the achievable read bandwidth and MFMA rate.
 * Return codes as in pirgpu.h (0 OK, 3 INVALID_ARGUMENT, 13 INTERNAL,
 * 14 UNAVAILABLE).
 */
#ifndef PIRGPU_PROBES_H_
#define PIRGPU_PROBES_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Streams `bytes` through a plain device-to-device copy kernel `repeats`
 * times and reports the average kernel time: the known-size kernel that
 * calibrates HBM byte counters. */
int pirg_copy_probe(size_t bytes, int repeats, double* ms);

/* Sweeps a plain read-only kernel (8 loads of 16 bytes in flight per thread)
 * `repeats` times over `bytes` within one launch and reports the time per
 * sweep: the achievable read bandwidth (HBM for large buffers, the caches
 * for small ones). */
int pirg_read_probe(size_t bytes, int repeats, double* ms);

/* Runs blocks x 256 threads of register-only v_mfma_i32_16x16x64_i8 (16
 * independent accumulators per wave, `iters` rounds) and reports the kernel
 * time and the int8 rate in TOP/s: the achievable MFMA rate. */
int pirg_mfma_probe(int blocks, int iters, double* ms, double* tops);

/* The same stream with random int8 operands (eight hashed operand pairs per
 * lane, rotated across the accumulators): the rate that products over random
 * data can approach, since the matrix cores' clock falls with the operands'
 * bit activity. */
int pirg_mfma_probe_random(int blocks, int iters, double* ms, double* tops);

#ifdef __cplusplus
}
#endif

#endif /* PIRGPU_PROBES_H_ */
