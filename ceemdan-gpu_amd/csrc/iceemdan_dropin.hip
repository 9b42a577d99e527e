// Drop-in iceemdanS / iceemdanD (cudaICEEMDAN.cuh:25-26) over the plan API.
// Reference semantics kept: Rilling stop rule capped at max_iter for the
// CEEMDAN stages and at 2000 for the noise-mode stages (cudaICEEMDAN.cu:851,
// :979), eps0 = noiseStrength, seed 1234 (cudaICEEMDAN.cu:740).  Arguments
// are validated (the reference checks none); d_x must be 0..N-1.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/cudaICEEMDAN.h"
#include "../../include/iceemdan_mi355x.h"
#include "icm_errors.h"
#include "icm_kernels.h"

namespace {

template <typename real>
double run(int dtype, size_t I, size_t N, size_t M, size_t max_iter, int* d_x, real* d_y,
           real* d_imfs, real eps0)
{
    auto fail = [](int st) -> double {
        std::fprintf(stderr, "iceemdan: %s\n", icm_status_string(st));
        return st < 0 ? (double)st : -1.0;
    };
    if (!d_x || !d_y || !d_imfs) return fail(ICM_ERR_INVALID_ARG);
    if (I < 1 || N < 4 || M < 2 || max_iter < 1 || I > 65535 || M > 64 || max_iter > 32000)
        return fail(ICM_ERR_INVALID_ARG);
    // d_x must be the sample numbers (cudaICEEMDAN.cu:423 compares them with p)
    int* d_err = nullptr;
    if (hipMalloc(&d_err, sizeof(int)) != hipSuccess) return fail(ICM_ERR_OUT_OF_MEMORY);
    int h_err = 0;
    bool ok = hipMemset(d_err, 0, sizeof(int)) == hipSuccess &&
              icm::launch_check_index(d_x, (int64_t)N, d_err, nullptr) == hipSuccess &&
              hipMemcpy(&h_err, d_err, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    (void)hipFree(d_err);
    if (!ok) return fail(ICM_ERR_HIP);
    if (h_err & ICM_DEVERR_INDEX) return fail(ICM_ERR_INDEX);

    icm_config c;
    icm_config_default(&c);
    int64_t len = (int64_t)N;
    c.dtype = dtype;
    c.num_channels = 1;
    c.lengths = &len;
    c.num_noise = (int)I;
    c.num_imfs = (int)M;
    c.stop_mode = ICM_STOP_RILLING;
    c.max_sifts = (int)max_iter;
    c.noise_max_sifts = 2000;
    c.noise_strength = (double)eps0;
    c.seed = 1234;
    c.use_graph = 0;
    icm_plan plan = nullptr;
    int st = icm_plan_create(&c, &plan);
    if (st) return fail(st);
    const void* sig[1] = {d_y};
    void* out[1] = {d_imfs};
    st = icm_forward(plan, sig, out, nullptr, nullptr);
    icm_timing t{};
    icm_last_timing(plan, &t);
    icm_plan_destroy(plan);
    if (st) return fail(st);
    return t.signal_ms * 1e-3;
}

}  // namespace

double iceemdanS(size_t numNoise, size_t SignalLength, size_t num_IMFs, size_t max_iter, int* d_x,
                 float* d_y, float* d_IMFs, float noiseStrength)
{
    return run<float>(ICM_DTYPE_F32, numNoise, SignalLength, num_IMFs, max_iter, d_x, d_y, d_IMFs,
                      noiseStrength);
}

double iceemdanD(size_t numNoise, size_t SignalLength, size_t num_IMFs, size_t max_iter, int* d_x,
                 double* d_y, double* d_IMFs, double noiseStrength)
{
    return run<double>(ICM_DTYPE_F64, numNoise, SignalLength, num_IMFs, max_iter, d_x, d_y, d_IMFs,
                       noiseStrength);
}

// C-linkage aliases for FFI callers (ctypes cannot bind C++ names)
extern "C" double icm_iceemdanS(size_t numNoise, size_t SignalLength, size_t num_IMFs,
                                size_t max_iter, int* d_x, float* d_y, float* d_IMFs,
                                float noiseStrength)
{
    return iceemdanS(numNoise, SignalLength, num_IMFs, max_iter, d_x, d_y, d_IMFs, noiseStrength);
}
extern "C" double icm_iceemdanD(size_t numNoise, size_t SignalLength, size_t num_IMFs,
                                size_t max_iter, int* d_x, double* d_y, double* d_IMFs,
                                double noiseStrength)
{
    return iceemdanD(numNoise, SignalLength, num_IMFs, max_iter, d_x, d_y, d_IMFs, noiseStrength);
}
