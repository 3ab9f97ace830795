// ppm_df.hip — linear operators along one axis of a strided volume (stage c).
//
//   out[o, j, i] = alpha * sum_t W[j, t] * v(o, start[j] + t, i)  (+ out[o, j, i] if accumulate)
//   v(o, m, i)   = sum_c chw[c] * in[o*so + m*sa + i*si + c]
//
// Any linear map along z from Fz to df (central difference, the finite-amplitude mean
// of compute_df_amp, the Giessibl integral, the transpose used by the adjoint) is one
// banded operator W/start built by the host in FP64; directional derivatives apply the
// same kernel along x or y.  HBM-bound: for the Fz -> df case one 4-byte read and one
// 4-byte write per output sample, the taps of neighbouring outputs overlapping in cache.
#include "ppm_common.hpp"

namespace ppm {
namespace {

struct AxisArgs {
    long long so, sa, si;
    float chw[4];
    int nch;
    int n_outer, n_axis_in, n_inner, taps, n_axis_out;
    float alpha;
    int accumulate;
};

__global__ __launch_bounds__(256) void k_axis_operator(const float* __restrict__ in, const float* __restrict__ W,
                                                       const int* __restrict__ start, AxisArgs p,
                                                       float* __restrict__ out) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per_o = (size_t)p.n_axis_out * p.n_inner;
    if (gid >= per_o * p.n_outer) return;
    const size_t o = gid / per_o;
    const size_t r = gid - o * per_o;
    const int j = (int)(r / p.n_inner);
    const int i = (int)(r - (size_t)j * p.n_inner);
    const float* base = in + o * p.so + (size_t)i * p.si;
    const float* w = W + (size_t)j * p.taps;
    const int s0 = start[j];
    float acc = 0.f;
    for (int t = 0; t < p.taps; ++t) {
        const int m = min(max(s0 + t, 0), p.n_axis_in - 1);   // index kept in range whatever start holds
        const float* q = base + (size_t)m * p.sa;
        float v = p.chw[0] * q[0];
        for (int c = 1; c < p.nch; ++c) v = fmaf(p.chw[c], q[c], v);
        acc = fmaf(w[t], v, acc);
    }
    acc *= p.alpha;
    if (p.accumulate) acc += out[gid];
    out[gid] = acc;
}

}  // namespace

int launch_axis_operator(const float* in, long long so, long long sa, long long si, const float chw[4], int nch,
                         int n_outer, int n_axis_in, int n_inner, const float* W, const int* start, int taps,
                         int n_axis_out, float alpha, int accumulate, float* out, hipStream_t stream) {
    AxisArgs p{};
    p.so = so; p.sa = sa; p.si = si;
    for (int c = 0; c < 4; ++c) p.chw[c] = chw[c];
    p.nch = nch;
    p.n_outer = n_outer; p.n_axis_in = n_axis_in; p.n_inner = n_inner; p.taps = taps; p.n_axis_out = n_axis_out;
    p.alpha = alpha; p.accumulate = accumulate;
    const size_t n = (size_t)n_outer * n_axis_out * n_inner;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_axis_operator, dim3(blocks), dim3(256), 0, stream, in, W, start, p, out);
    return check_launch("axis_operator");
}

}  // namespace ppm
