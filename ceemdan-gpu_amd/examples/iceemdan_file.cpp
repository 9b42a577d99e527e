// Command-line driver: decompose raw float32 signal files.
//
// The counterpart of the reference's sample_binary_file.cu (read a raw
// little-endian float32 file, call iceemdanS, write [M][N] float32 modes)
// and sample_synthetic_signal.cu (the two-tone test signal), extended to
// several channels in one batched plan through the C ABI.
//
//   iceemdan_file [options] OUT IN [IN ...]
//   iceemdan_file [options] --two-tone OUT
//
// One input: OUT is the [M][N] float32 file.  Several inputs: OUT is a prefix
// and channel c is written to OUT.c.bin.
//
// options:
//   --noise I        realizations (default 200, sample_binary_file.cu:39)
//   --imfs M         output rows (default 12, :40)
//   --max-iter K     Rilling cap per stage (default 300, :41)
//   --eps0 E         noise strength (default 0.2, :42)
//   --fixed S        FIXED rule with S sifts per stage and per noise mode
//                    (default: the reference's Rilling rule)
//   --double         compute in float64 (input and output files stay float32)
//   --two-tone       use the reference's synthetic signal (N = 1000,
//                    sample_synthetic_signal.cu:48-55) instead of input files
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cudaICEEMDAN.h"
#include "../../include/iceemdan_mi355x.h"

namespace {

struct Options {
    int noise = 200, imfs = 12, max_iter = 300, fixed = 0;
    double eps0 = 0.2;
    bool dbl = false, two_tone = false;
    std::string out;
    std::vector<std::string> inputs;
};

void usage()
{
    std::fprintf(stderr,
                 "usage: iceemdan_file [--noise I] [--imfs M] [--max-iter K] [--eps0 E] "
                 "[--fixed S] [--double] OUT IN [IN ...]\n"
                 "       iceemdan_file [options] --two-tone OUT\n");
}

bool parse(int argc, char** argv, Options& o)
{
    std::vector<std::string> pos;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char* what) -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "%s needs a value\n", what);
                return nullptr;
            }
            return argv[++i];
        };
        const char* v = nullptr;
        if (a == "--noise") { if (!(v = next("--noise"))) return false; o.noise = std::atoi(v); }
        else if (a == "--imfs") { if (!(v = next("--imfs"))) return false; o.imfs = std::atoi(v); }
        else if (a == "--max-iter") { if (!(v = next("--max-iter"))) return false; o.max_iter = std::atoi(v); }
        else if (a == "--eps0") { if (!(v = next("--eps0"))) return false; o.eps0 = std::atof(v); }
        else if (a == "--fixed") { if (!(v = next("--fixed"))) return false; o.fixed = std::atoi(v); }
        else if (a == "--double") o.dbl = true;
        else if (a == "--two-tone") o.two_tone = true;
        else if (a == "-h" || a == "--help") return false;
        else pos.push_back(a);
    }
    if (pos.empty()) return false;
    o.out = pos[0];
    o.inputs.assign(pos.begin() + 1, pos.end());
    if (o.two_tone != o.inputs.empty()) return false;
    return o.noise > 0 && o.imfs >= 2 && o.max_iter > 0 && o.fixed >= 0;
}

bool read_f32(const std::string& path, std::vector<float>& x)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path.c_str());
        return false;
    }
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    x.resize((size_t)bytes / sizeof(float));
    const size_t got = std::fread(x.data(), sizeof(float), x.size(), f);
    std::fclose(f);
    if (got != x.size() || x.size() < 4) {
        std::fprintf(stderr, "short or empty file %s\n", path.c_str());
        return false;
    }
    return true;
}

bool write_f32(const std::string& path, const float* p, size_t n)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        return false;
    }
    const size_t put = std::fwrite(p, sizeof(float), n, f);
    std::fclose(f);
    return put == n;
}

// sample_synthetic_signal.cu:48-55: a 0.065 cycles/sample tone plus a
// 0.255 cycles/sample burst on samples 500..748, with that file's value of pi
std::vector<float> two_tone()
{
    const double pi = 3.1415926;
    std::vector<float> y(1000);
    for (int i = 0; i < 1000; ++i) {
        double v = std::sin(2 * pi * 0.065 * i);
        if (i >= 500 && i <= 748) v += std::sin(2 * pi * 0.255 * (i - 500));
        y[i] = (float)v;
    }
    return y;
}

#define HIP_OK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));           \
            return 1;                                                              \
        }                                                                          \
    } while (0)

// One channel through the drop-in entry point (the reference's call).
template <typename real>
int run_dropin(const Options& o, const std::vector<float>& y, std::vector<float>& modes)
{
    const size_t N = y.size(), M = (size_t)o.imfs;
    std::vector<int> hx(N);
    for (size_t i = 0; i < N; ++i) hx[i] = (int)i;
    std::vector<real> hy(y.begin(), y.end());
    int* d_x = nullptr;
    real *d_y = nullptr, *d_imfs = nullptr;
    HIP_OK(hipMalloc(&d_x, N * sizeof(int)));
    HIP_OK(hipMalloc(&d_y, N * sizeof(real)));
    HIP_OK(hipMalloc(&d_imfs, M * N * sizeof(real)));
    HIP_OK(hipMemcpy(d_x, hx.data(), N * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_y, hy.data(), N * sizeof(real), hipMemcpyHostToDevice));
    double t;
    if constexpr (sizeof(real) == 4)
        t = iceemdanS((size_t)o.noise, N, M, (size_t)o.max_iter, d_x, d_y, d_imfs, (float)o.eps0);
    else
        t = iceemdanD((size_t)o.noise, N, M, (size_t)o.max_iter, d_x, d_y, d_imfs, o.eps0);
    if (t < 0) return 1;
    std::printf("Execution time: %f\n", t);
    std::vector<real> out(M * N);
    HIP_OK(hipMemcpy(out.data(), d_imfs, M * N * sizeof(real), hipMemcpyDeviceToHost));
    modes.assign(out.begin(), out.end());
    (void)hipFree(d_x);
    (void)hipFree(d_y);
    (void)hipFree(d_imfs);
    return 0;
}

// Many channels in one plan through the C ABI.
template <typename real>
int run_plan(const Options& o, const std::vector<std::vector<float>>& ys,
             std::vector<std::vector<float>>& modes)
{
    const int C = (int)ys.size(), M = o.imfs;
    icm_config cfg;
    icm_config_default(&cfg);
    std::vector<int64_t> lens(C);
    for (int c = 0; c < C; ++c) lens[c] = (int64_t)ys[c].size();
    cfg.dtype = sizeof(real) == 8 ? ICM_DTYPE_F64 : ICM_DTYPE_F32;
    cfg.num_channels = C;
    cfg.lengths = lens.data();
    cfg.num_noise = o.noise;
    cfg.num_imfs = M;
    cfg.noise_strength = o.eps0;
    if (o.fixed > 0) {
        cfg.stop_mode = ICM_STOP_FIXED;
        cfg.max_sifts = cfg.noise_max_sifts = o.fixed;
    } else {
        cfg.stop_mode = ICM_STOP_RILLING;
        cfg.max_sifts = o.max_iter;
        cfg.noise_max_sifts = 2000;
    }
    icm_plan plan = nullptr;
    int rc = icm_plan_create(&cfg, &plan);
    if (rc) {
        std::fprintf(stderr, "icm_plan_create: %s\n", icm_status_string(rc));
        return 1;
    }
    std::vector<void*> d_in(C), d_out(C);
    for (int c = 0; c < C; ++c) {
        std::vector<real> h(ys[c].begin(), ys[c].end());
        HIP_OK(hipMalloc(&d_in[c], h.size() * sizeof(real)));
        HIP_OK(hipMalloc(&d_out[c], (size_t)M * h.size() * sizeof(real)));
        HIP_OK(hipMemcpy(d_in[c], h.data(), h.size() * sizeof(real), hipMemcpyHostToDevice));
    }
    rc = icm_forward(plan, (const void* const*)d_in.data(), d_out.data(), nullptr, nullptr);
    if (rc) {
        std::fprintf(stderr, "icm_forward: %s %s\n", icm_status_string(rc), icm_last_error());
        return 1;
    }
    icm_timing tm;
    icm_last_timing(plan, &tm);
    std::printf("Execution time: %f (noise modes %f, total %f)\n", tm.signal_ms * 1e-3,
                tm.noise_ms * 1e-3, tm.total_ms * 1e-3);
    modes.resize(C);
    for (int c = 0; c < C; ++c) {
        std::vector<real> h((size_t)M * ys[c].size());
        HIP_OK(hipMemcpy(h.data(), d_out[c], h.size() * sizeof(real), hipMemcpyDeviceToHost));
        modes[c].assign(h.begin(), h.end());
        (void)hipFree(d_in[c]);
        (void)hipFree(d_out[c]);
    }
    icm_plan_destroy(plan);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    if (!parse(argc, argv, o)) {
        usage();
        return 2;
    }
    std::vector<std::vector<float>> ys;
    if (o.two_tone) {
        ys.push_back(two_tone());
    } else {
        for (const auto& p : o.inputs) {
            ys.emplace_back();
            if (!read_f32(p, ys.back())) return 1;
        }
    }
    std::vector<std::vector<float>> modes;
    int rc;
    if (ys.size() == 1 && o.fixed == 0) {
        modes.resize(1);
        rc = o.dbl ? run_dropin<double>(o, ys[0], modes[0]) : run_dropin<float>(o, ys[0], modes[0]);
    } else {
        rc = o.dbl ? run_plan<double>(o, ys, modes) : run_plan<float>(o, ys, modes);
    }
    if (rc) return rc;
    if (modes.size() == 1)
        return write_f32(o.out, modes[0].data(), modes[0].size()) ? 0 : 1;
    for (size_t c = 0; c < modes.size(); ++c)
        if (!write_f32(o.out + "." + std::to_string(c) + ".bin", modes[c].data(), modes[c].size()))
            return 1;
    return 0;
}
