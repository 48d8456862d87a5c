// Seeded synthetic workloads (host and device). Every draw is a pure function
// of (seed, stream, a, b) through a splitmix64 hash chain, so host and device
// produce the same integer draws; lognormal/gamma transforms are evaluated in
// double precision and rounded to fp32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ddrs_hip.h"

extern "C" void ddrs_internal_set_error(const char* msg);

namespace {

__host__ __device__ inline uint64_t mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t hash4(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b) {
    return mix(mix(mix(seed ^ (stream * 0xD1B54A32D192ED03ull)) ^ a) ^ (b * 0x9E3779B97F4A7C15ull));
}
// uniform in (0, 1)
__host__ __device__ inline double u01(uint64_t h) {
    return (static_cast<double>(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

enum : uint64_t { ST_TREE = 1, ST_LEN, ST_SLOPE, ST_N, ST_QSP, ST_Z1, ST_Z2, ST_STORM, ST_G1, ST_G2 };

// q'[t][r] for one reach r over all t (storm state is sequential in t)
__host__ __device__ inline void forcing_reach(int n, int T, uint64_t seed, int storms, int r,
                                              float* q) {
    const double decay = exp(-1.0 / 24.0);
    double m = 0.0;
    for (int t = 0; t < T; ++t) {
        const double u1 = u01(hash4(seed, ST_Z1, t, r));
        const double u2 = u01(hash4(seed, ST_Z2, t, r));
        const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        double v = exp(log(0.5) + 1.0 * z);
        if (storms) {
            m *= decay;
            if (u01(hash4(seed, ST_STORM, t, r)) < 1.0 / 72.0) {
                const double ga = u01(hash4(seed, ST_G1, t, r));
                const double gb = u01(hash4(seed, ST_G2, t, r));
                m += -2.0 * (log(ga) + log(gb));  // Gamma(shape 2, scale 2)
            }
            v *= (1.0 + m);
        }
        q[static_cast<size_t>(t) * n + r] = static_cast<float>(v);
    }
}

__global__ void k_forcing(int n, int T, uint64_t seed, int storms, float* q) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) forcing_reach(n, T, seed, storms, r, q);
}

}  // namespace

extern "C" {

ddrs_status ddrs_gen_tree(int32_t n, int32_t window, uint64_t seed, int32_t* h_rows,
                          int32_t* h_cols) {
    if (n < 1 || window < 1 || (n > 1 && (!h_rows || !h_cols))) {
        ddrs_internal_set_error("ddrs_gen_tree: bad arguments");
        return DDRS_ERR_INVALID;
    }
    for (int i = 0; i + 1 < n; ++i) {
        const int hi = std::min(i + window, n - 1);
        int d = i + 1 + static_cast<int>(u01(hash4(seed, ST_TREE, i, 0)) * (hi - i));
        d = std::min(d, hi);
        h_rows[i] = d;
        h_cols[i] = i;
    }
    return DDRS_OK;
}

ddrs_status ddrs_gen_attributes(int32_t n, uint64_t seed, float* h_length, float* h_slope,
                                float* h_n, float* h_q_spatial) {
    if (n < 0) {
        ddrs_internal_set_error("ddrs_gen_attributes: bad arguments");
        return DDRS_ERR_INVALID;
    }
    const double ls0 = std::log(1e-4), ls1 = std::log(5e-2);
    for (int i = 0; i < n; ++i) {
        if (h_length) h_length[i] = static_cast<float>(1000.0 + 9000.0 * u01(hash4(seed, ST_LEN, i, 0)));
        if (h_slope) h_slope[i] = static_cast<float>(std::exp(ls0 + (ls1 - ls0) * u01(hash4(seed, ST_SLOPE, i, 0))));
        if (h_n) h_n[i] = static_cast<float>(0.015 + 0.235 * u01(hash4(seed, ST_N, i, 0)));
        if (h_q_spatial) h_q_spatial[i] = static_cast<float>(u01(hash4(seed, ST_QSP, i, 0)));
    }
    return DDRS_OK;
}

ddrs_status ddrs_gen_forcing(int32_t n, int32_t T, uint64_t seed, int32_t storms,
                             float* h_q_prime) {
    if (n < 0 || T < 0 || !h_q_prime) {
        ddrs_internal_set_error("ddrs_gen_forcing: bad arguments");
        return DDRS_ERR_INVALID;
    }
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned k = 0; k < nt; ++k)
        th.emplace_back([=] {
            for (int r = static_cast<int>(k); r < n; r += static_cast<int>(nt))
                forcing_reach(n, T, seed, storms, r, h_q_prime);
        });
    for (auto& t : th) t.join();
    return DDRS_OK;
}

ddrs_status ddrs_gen_forcing_device(int32_t n, int32_t T, uint64_t seed, int32_t storms,
                                    float* d_q_prime, void* stream) {
    if (n < 0 || T < 0 || !d_q_prime) {
        ddrs_internal_set_error("ddrs_gen_forcing_device: bad arguments");
        return DDRS_ERR_INVALID;
    }
    if (n == 0) return DDRS_OK;
    hipLaunchKernelGGL(k_forcing, dim3((n + 255) / 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), n, T, seed, storms, d_q_prime);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ddrs_internal_set_error(hipGetErrorString(e));
        return DDRS_ERR_HIP;
    }
    return DDRS_OK;
}

}  // extern "C"
