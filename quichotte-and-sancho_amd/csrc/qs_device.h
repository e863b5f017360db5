// qs_device.h -- device helpers shared by the gfx950 kernels of the SCF path.
//
// Complex numbers are double2 (x = real, y = imaginary) in FP64.  Wavefronts
// are 64 lanes wide; block reductions assume blockDim.x is a multiple of 64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define QS_WAVE 64

struct cplx { double x, y; };

__device__ __forceinline__ double2 cmk(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// acc += a * b
__device__ __forceinline__ void cfma(double2 &acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
// acc -= a * b
__device__ __forceinline__ void cfms(double2 &acc, double2 a, double2 b) {
    acc.x = fma(-a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(-a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
}
// acc += a * conj(b)
__device__ __forceinline__ void cfmac(double2 &acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(-a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
__device__ __forceinline__ double2 cinv(double2 a) {
    const double r = 1.0 / (a.x * a.x + a.y * a.y);
    return make_double2(a.x * r, -a.y * r);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = QS_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, QS_WAVE);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int o = QS_WAVE / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, QS_WAVE));
    return v;
}

// Block-wide sum; every thread gets the result.  scratch: >= 16 doubles of LDS.
__device__ __forceinline__ double block_sum(double v, double *scratch) {
    v = wave_sum(v);
    int w = threadIdx.x / QS_WAVE, nw = blockDim.x / QS_WAVE;
    __syncthreads();
    if ((threadIdx.x & (QS_WAVE - 1)) == 0) scratch[w] = v;
    __syncthreads();
    double s = 0;
    for (int i = 0; i < nw; i++) s += scratch[i];
    return s;
}
__device__ __forceinline__ double block_max(double v, double *scratch) {
    v = wave_max(v);
    int w = threadIdx.x / QS_WAVE, nw = blockDim.x / QS_WAVE;
    __syncthreads();
    if ((threadIdx.x & (QS_WAVE - 1)) == 0) scratch[w] = v;
    __syncthreads();
    double s = scratch[0];
    for (int i = 1; i < nw; i++) s = fmax(s, scratch[i]);
    return s;
}

// Largest 64-bit key of the wavefront, on every lane: DPP shifts within rows of
// 16 lanes, then row broadcasts 15 and 31 (gfx9 DPP), then lane 63.
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t k) {
    auto step = [&](auto ctrl_c, auto rmask_c) {
        constexpr int ctrl = decltype(ctrl_c)::value, rmask = decltype(rmask_c)::value;
        const int lo = (int)(uint32_t)k, hi = (int)(uint32_t)(k >> 32);
        const uint32_t olo = (uint32_t)__builtin_amdgcn_update_dpp(lo, lo, ctrl, rmask, 0xf, false);
        const uint32_t ohi = (uint32_t)__builtin_amdgcn_update_dpp(hi, hi, ctrl, rmask, 0xf, false);
        const uint64_t o = ((uint64_t)ohi << 32) | olo;
        k = o > k ? o : k;
    };
    step(std::integral_constant<int, 0x111>{}, std::integral_constant<int, 0xf>{});   // row_shr:1
    step(std::integral_constant<int, 0x112>{}, std::integral_constant<int, 0xf>{});   // row_shr:2
    step(std::integral_constant<int, 0x114>{}, std::integral_constant<int, 0xf>{});   // row_shr:4
    step(std::integral_constant<int, 0x118>{}, std::integral_constant<int, 0xf>{});   // row_shr:8
    step(std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xa>{});   // row_bcast:15
    step(std::integral_constant<int, 0x143>{}, std::integral_constant<int, 0xc>{});   // row_bcast:31
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(k >> 32), 63);
    return ((uint64_t)hi << 32) | lo;
}

// The same over lanes 0..31 only (lanes 32..63 must hold 0), on every lane: four
// DPP shifts within rows of 16, then row broadcast 15, then lane 31.  The shifts
// fill from outside their row with 0 (bound_ctrl), the identity of the max, so
// they need no copy of the old value.
__device__ __forceinline__ uint64_t wave_max_u64_lo32(uint64_t k) {
    auto step = [&](auto ctrl_c, auto rmask_c, auto bc_c) {
        constexpr int ctrl = decltype(ctrl_c)::value, rmask = decltype(rmask_c)::value;
        constexpr bool bc = decltype(bc_c)::value;
        const int lo = (int)(uint32_t)k, hi = (int)(uint32_t)(k >> 32);
        const uint32_t olo = (uint32_t)__builtin_amdgcn_update_dpp(bc ? 0 : lo, lo, ctrl, rmask, 0xf, bc);
        const uint32_t ohi = (uint32_t)__builtin_amdgcn_update_dpp(bc ? 0 : hi, hi, ctrl, rmask, 0xf, bc);
        const uint64_t o = ((uint64_t)ohi << 32) | olo;
        k = o > k ? o : k;
    };
    using T = std::true_type;
    using F = std::false_type;
    step(std::integral_constant<int, 0x111>{}, std::integral_constant<int, 0xf>{}, T{});   // row_shr:1
    step(std::integral_constant<int, 0x112>{}, std::integral_constant<int, 0xf>{}, T{});   // row_shr:2
    step(std::integral_constant<int, 0x114>{}, std::integral_constant<int, 0xf>{}, T{});   // row_shr:4
    step(std::integral_constant<int, 0x118>{}, std::integral_constant<int, 0xf>{}, T{});   // row_shr:8
    step(std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xa>{}, F{});   // row_bcast:15
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k, 31);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(k >> 32), 31);
    return ((uint64_t)hi << 32) | lo;
}

// Fermi-Dirac as the reference evaluates it (src/green_functions.f90:74-84):
// f = 1/(exp(min(hbar*(w - mu + V/hbar)*beta, 700)) + 1), hbar = 1.
__device__ __host__ __forceinline__ double qs_fermi(double w, double V, double mu, double beta) {
    double e = (w - mu + V) * beta;
    if (e > 700.0) e = 700.0;
    return 1.0 / (exp(e) + 1.0);
}

// Site-pair layout of G and Sigma2 on the active block (the FFT columns and the
// Sigma2 read by the Dyson kernels).  Per unordered active site pair (I <= J),
// pair index sig_pair(I, J) (row-major upper triangle), 12 complex values:
//   [0, 4)  retarded block (I, J), element 2 s_i + s_j
//   [4, 8)  lesser   block (I, J), element 2 s_i + s_j
//   [8, 12) retarded block (J, I), element 2 s_j + s_i (row spin first)
// The lesser block (J, I) is not stored: G^< and Sigma^< are anti-Hermitian,
// L(J, I) = -L(I, J)^dagger.  For I = J slots 8-11 repeat slots 0-3.
__host__ __device__ inline int sig_pair(int I, int J, int nAs) { return I * nAs - (I * (I - 1)) / 2 + (J - I); }
// retarded element (a, b) of the active block (a, b: active spin-orbital indices)
__host__ __device__ inline int sig_r_index(int a, int b, int nAs) {
    const int I = a >> 1, J = b >> 1;
    return I <= J ? 12 * sig_pair(I, J, nAs) + ((a & 1) << 1) + (b & 1)
                  : 12 * sig_pair(J, I, nAs) + 8 + ((a & 1) << 1) + (b & 1);
}
// lesser element (a, b): the stored index, and whether the element is -conj of it
struct SigL {
    int idx;
    bool flip;
};
__host__ __device__ inline SigL sig_l_at(int a, int b, int nAs) {
    const int I = a >> 1, J = b >> 1;
    SigL r;
    r.flip = I > J;
    r.idx = I <= J ? 12 * sig_pair(I, J, nAs) + 4 + ((a & 1) << 1) + (b & 1)
                   : 12 * sig_pair(J, I, nAs) + 4 + ((b & 1) << 1) + (a & 1);
    return r;
}
__device__ __forceinline__ double2 sig_l_value(double2 v, bool flip) { return flip ? make_double2(-v.x, v.y) : v; }
