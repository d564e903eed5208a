// eu3d_mlp_internal.h -- what the solver's chemistry needs from the network
// object (eu3d_mlp.hip): the padded input/output buffers, the normalization
// table and a forward pass on a given stream.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

struct eu3d_mlp;

namespace eu3d_mlp_internal {

constexpr int kIn = 32;  // padded input width (features beyond n_in are 0)
constexpr int kOut = 32; // padded output width

typedef unsigned short bf16;

__host__ __device__ __forceinline__ bf16 f2bf(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) // inf / nan: truncate, keep nan quiet
        return (bf16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even
    return (bf16)(u >> 16);
}

__host__ __device__ __forceinline__ float bf2f(bf16 b)
{
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// norm layout: in_mean[kIn], in_std[kIn], out_mean[kOut], out_std[kOut]
struct View {
    bf16 *X0;          // [max_batch x kIn] normalized inputs
    const float *Y;    // [max_batch x kOut] normalized outputs
    const float *norm; // device
    int n_in, n_out, max_batch;
};

View view(eu3d_mlp *mlp);
// Inference forward of the first B rows (B a multiple of 64, <= max_batch)
// on the given stream.  Returns an EU3D_* status.
int forward(eu3d_mlp *mlp, hipStream_t stream, int B);

} // namespace eu3d_mlp_internal
