/*
 * Drop-in host entry points of the MI355X ICEEMDAN library.
 *
 * Same names, argument order and meaning as the reference header
 * cudaICEEMDAN.cuh:25-26 (definitions cudaICEEMDAN.cu:619 and :1340), with C++
 * linkage like the reference.  Callers: sample_synthetic_signal.cu:70,
 * sample_binary_file.cu:80, MatlabMEXCUDA/iceemdanCUDA_MexFun.cu:72.
 *
 *   numNoise       I, noise realizations
 *   SignalLength   N, samples
 *   num_IMFs       M, rows of d_IMFs (M-1 modes + residue)
 *   max_iter       sifting cap per stage (Rilling stop rule)
 *   d_x            device int[N], must hold 0..N-1 (the interval search uses
 *                  the sample number, cudaICEEMDAN.cu:418-436)
 *   d_y            device real[N] input, read only
 *   d_IMFs         device real[M][N] output, fully overwritten, in units of
 *                  x / std(x)
 *   noiseStrength  eps0
 * Returns the seconds spent in the CEEMDAN stages (the reference's timer,
 * cudaICEEMDAN.cu:1002-1297), or a negative value on error (the status string
 * is printed to stderr).  The noise-mode stages use the Rilling rule with the
 * reference's cap of 2000 passes (cudaICEEMDAN.cu:851); the white noise is
 * Philox4x64-10 with seed 1234 (the reference uses cuRAND XORWOW, seed 1234).
 */
#ifndef ICM_CUDAICEEMDAN_H
#define ICM_CUDAICEEMDAN_H
#include <stddef.h>

double iceemdanS(size_t numNoise, size_t SignalLength, size_t num_IMFs, size_t max_iter, int* d_x,
                 float* d_y, float* d_IMFs, float noiseStrength);
double iceemdanD(size_t numNoise, size_t SignalLength, size_t num_IMFs, size_t max_iter, int* d_x,
                 double* d_y, double* d_IMFs, double noiseStrength);

#endif
