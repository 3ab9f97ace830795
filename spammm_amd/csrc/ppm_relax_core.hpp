// ppm_relax_core.hpp — device code of the probe relaxation shared by the plain and the
// batched kernels.  Translation units that include it are built with -ffp-contract=off: every
// FP32 operation is evaluated exactly as written (see ppm_relax.hip for the semantics and
// the reference kernels they follow).
#pragma once
#pragma clang fp contract(off)
#include <type_traits>

#include "ppm_common.hpp"

namespace ppm {
namespace rc {

constexpr float kFTDEC = 0.5f, kFTINC = 1.1f, kFDAMP = 0.99f, kF2SAFE = 1e-8f;

// A grid as the relaxation sees it.  Only the z planes [klo, klo + nzw) are stored (the
// whole grid when klo = 0, nzw = nz); a sample that needs a plane outside them sets *flag and
// reads clamped planes instead (its result is then not used: the caller recomputes with the
// whole grid).
struct GridDev {
    const float4* __restrict__ data;
    int nx, ny, nz;
    float da[4], db[4], dc[4];
    int klo, nzw;
    int* flag;
};

// The cached cell is named by the floor values of the three axes (axis_frac): the corner indices
// are a function of them (axis_index), so they are only computed when one changes.
struct Cell {
    float fi, fj, fk;  // floors of the cached cell; -2 = empty
    float4 c[8];       // corners 000,100,010,110,001,101,011,111
    __device__ void reset() { fi = fj = fk = -2.0f; }
};

// REPEAT addressing of one axis: the interpolation weight a and the floor fl of the texel
// coordinate (in [-1, n - 1] for a finite coordinate).
__device__ __forceinline__ float axis_frac(float s, int n, float& fl) {
    const float u = (s - floorf(s)) * (float)n;
    const float t = u - 0.5f;
    fl = floorf(t);
    return t - fl;
}

// The two texel indices of the cell with floor fl, REPEAT-wrapped.  fl <= n - 1 for every finite
// coordinate (the fraction is below 1, or rounds to 1 and gives t = n - 0.5), and fmaxf maps a
// NaN floor (non-finite coordinate) to -1, so i is in [-1, n - 1] and one wrap per index suffices.
__device__ __forceinline__ void axis_index(float fl, int n, int& i0, int& i1) {
    const int i = (int)fmaxf(fl, -1.0f);
    i0 = i < 0 ? n - 1 : i;
    i1 = i + 1 > n - 1 ? 0 : i + 1;
}

// Stored planes hold at most 4 GiB: corner offsets fit 32 bits.
__host__ __device__ inline bool small_grid(const GridDev& g) {
    return (unsigned long long)g.nx * (unsigned long long)g.ny * (unsigned long long)g.nzw * 16ull <=
           0xFFFFFFFFull;
}

// Element offsets of the 4 corner rows (k0,j0), (k0,j1), (k1,j0), (k1,j1) of the cell with floors
// (fa, fb, fc) in the stored planes, and its two x indices.  A cell outside the stored planes sets
// *g.flag (from lane `flag_lane`) and is clamped into them.  O32: 32-bit offsets (small_grid).
template <bool O32>
__device__ __forceinline__ void cell_rows(const GridDev& g, float fa, float fb, float fc, bool flag_lane, int& i0,
                                          int& i1, typename std::conditional<O32, unsigned, size_t>::type (&r)[4]) {
    using T = typename std::conditional<O32, unsigned, size_t>::type;
    int j0, j1, k0, k1;
    axis_index(fa, g.nx, i0, i1);
    axis_index(fb, g.ny, j0, j1);
    axis_index(fc, g.nz, k0, k1);
    k0 -= g.klo;
    k1 -= g.klo;
    if ((unsigned)k0 >= (unsigned)g.nzw || (unsigned)k1 >= (unsigned)g.nzw) {
        if (g.flag && flag_lane) atomicOr(g.flag, 1);
        k0 = min(max(k0, 0), g.nzw - 1);
        k1 = min(max(k1, 0), g.nzw - 1);
    }
    const T sx = (T)g.nx, sxy = (T)g.nx * (T)g.ny;
    const T z0 = (T)k0 * sxy, z1 = (T)k1 * sxy, y0 = (T)j0 * sx, y1 = (T)j1 * sx;
    r[0] = z0 + y0; r[1] = z0 + y1; r[2] = z1 + y0; r[3] = z1 + y1;
}

__device__ __forceinline__ float4 lerp8(const Cell& C, float a, float b, float c) {
    const float ma = 1.0f - a, mb = 1.0f - b, mc = 1.0f - c;
    float4 o;
#define PPM_TRI(q)                                                                                   \
    {                                                                                               \
        const float lo = (C.c[0].q * ma + C.c[1].q * a) * mb + (C.c[2].q * ma + C.c[3].q * a) * b;   \
        const float hi = (C.c[4].q * ma + C.c[5].q * a) * mb + (C.c[6].q * ma + C.c[7].q * a) * b;   \
        o.q = lo * mc + hi * c;                                                                      \
    }
    PPM_TRI(x) PPM_TRI(y) PPM_TRI(z) PPM_TRI(w)
#undef PPM_TRI
    return o;
}

// Texture coordinates of a world point.  AX: the grid axes are the world axes (the off-diagonal
// entries of dinv are zero), so px*d[0] + py*0 + pz*0 + d[3] = px*d[0] + d[3] for finite
// coordinates (the dropped terms are signed zeros) and the two multiplies are skipped.
template <bool AX>
__device__ __forceinline__ void tex_coords(const GridDev& g, float px, float py, float pz, float& sa, float& sb,
                                           float& sc) {
    if (AX) {
        sa = px * g.da[0] + g.da[3];
        sb = py * g.db[1] + g.db[3];
        sc = pz * g.dc[2] + g.dc[3];
    } else {
        sa = px * g.da[0] + py * g.da[1] + pz * g.da[2] + g.da[3];
        sb = px * g.db[0] + py * g.db[1] + pz * g.db[2] + g.db[3];
        sc = px * g.dc[0] + py * g.dc[1] + pz * g.dc[2] + g.dc[3];
    }
}

// FAST: the grid axes are the world axes (tex_coords<true>) and the stored planes are a small_grid
// (32-bit corner offsets from the uniform grid base).
template <bool FAST = false>
__device__ __forceinline__ float4 sample(const GridDev& g, Cell& C, float px, float py, float pz) {
    float sa, sb, sc;
    tex_coords<FAST>(g, px, py, pz, sa, sb, sc);
    float fa, fb, fc;
    const float a = axis_frac(sa, g.nx, fa), b = axis_frac(sb, g.ny, fb), c = axis_frac(sc, g.nz, fc);
    if (fa != C.fi || fb != C.fj || fc != C.fk) {
        C.fi = fa; C.fj = fb; C.fk = fc;
        int i0, i1;
        typename std::conditional<FAST, unsigned, size_t>::type r[4];
        cell_rows<FAST>(g, fa, fb, fc, true, i0, i1, r);
        if constexpr (FAST) {
            const char* base = reinterpret_cast<const char*>(g.data);
#pragma unroll
            for (int m = 0; m < 8; ++m)
                C.c[m] = *reinterpret_cast<const float4*>(base + (r[m >> 1] + (unsigned)((m & 1) ? i1 : i0)) * 16u);
        } else {
#pragma unroll
            for (int m = 0; m < 8; ++m) C.c[m] = g.data[r[m >> 1] + (size_t)((m & 1) ? i1 : i0)];
        }
    }
    return lerp8(C, a, b, c);
}

// The same sample shared by a group of G lanes (G = 2 or 4) that relax one scan point: lane
// `sub` caches and interpolates only channels [sub*4/G, (sub+1)*4/G) of the 8 corners, with
// the per-channel formula of lerp8 (so every channel is bit-identical to sample()), and the
// group assembles the float4 with shuffles.  The group's lanes hold the same position, so
// they take the same branches and every shuffle reads an active lane.
template <int G>
struct GroupCell {
    float fi, fj, fk;
    float c[4 / G][8];
    __device__ void reset() { fi = fj = fk = -2.0f; }
};

template <int G, bool FAST = false>
__device__ __forceinline__ float4 sample_group(const GridDev& g, GroupCell<G>& C, int sub, float px, float py,
                                               float pz) {
    constexpr int NC = 4 / G;
    float sa, sb, sc;
    tex_coords<FAST>(g, px, py, pz, sa, sb, sc);
    float fa, fb, fc;
    const float a = axis_frac(sa, g.nx, fa), b = axis_frac(sb, g.ny, fb), c = axis_frac(sc, g.nz, fc);
    if (fa != C.fi || fb != C.fj || fc != C.fk) {
        C.fi = fa; C.fj = fb; C.fk = fc;
        int i0, i1;
        typename std::conditional<FAST, unsigned, size_t>::type r[4];
        cell_rows<FAST>(g, fa, fb, fc, sub == 0, i0, i1, r);
        if constexpr (FAST) {
            const char* base = reinterpret_cast<const char*>(g.data);
            const unsigned ch = (unsigned)(sub * NC * 4);       // this lane's channels within a texel
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const char* t = base + ((r[m >> 1] + (unsigned)((m & 1) ? i1 : i0)) * 16u + ch);
                if constexpr (NC == 2) {
                    const float2 v = *reinterpret_cast<const float2*>(t);
                    C.c[0][m] = v.x;
                    C.c[1][m] = v.y;
                } else {
#pragma unroll
                    for (int q = 0; q < NC; ++q) C.c[q][m] = reinterpret_cast<const float*>(t)[q];
                }
            }
        } else {
            const float* base = reinterpret_cast<const float*>(g.data);
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const size_t idx = r[m >> 1] + (size_t)((m & 1) ? i1 : i0);
#pragma unroll
                for (int q = 0; q < NC; ++q) C.c[q][m] = base[4 * idx + sub * NC + q];
            }
        }
    }
    const float ma = 1.0f - a, mb = 1.0f - b, mc = 1.0f - c;
    float v[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        const float* f = C.c[q];
        const float lo = (f[0] * ma + f[1] * a) * mb + (f[2] * ma + f[3] * a) * b;
        const float hi = (f[4] * ma + f[5] * a) * mb + (f[6] * ma + f[7] * a) * b;
        v[q] = lo * mc + hi * c;
    }
    float o[4];
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) o[ch] = __shfl(v[ch % NC], ch / NC, G);
    return make_float4(o[0], o[1], o[2], o[3]);
}

__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 rotMat(f3 v, f3 a, f3 b, f3 c) { return mk3(dot3(v, a), dot3(v, b), dot3(v, c)); }
__device__ __forceinline__ f3 rotMatT(f3 v, f3 a, f3 b, f3 c) {
    return mk3(a.x * v.x + b.x * v.y + c.x * v.z, a.y * v.x + b.y * v.y + c.y * v.z,
               a.z * v.x + b.z * v.y + c.z * v.z);
}
__device__ __forceinline__ f3 tipForce(f3 d, const float* k, const float* d0) {
    float r = sqrtf(dot3(d, d));
    r = fmaxf(r, 1e-10f);
    const float s = (k[3] * (r - d0[3])) / r;
    return mk3((d.x - d0[0]) * k[0] + d.x * s, (d.y - d0[1]) * k[1] + d.y * s, (d.z - d0[2]) * k[2] + d.z * s);
}
// update_FIRE: returns the new velocity, updates dt and damp.
__device__ __forceinline__ f3 fire(f3 f, f3 v, float& dt, float& damp, float dtmin, float dtmax, float damp0) {
    const float ff = dot3(f, f), vv = dot3(v, v), vf = dot3(v, f);
    if (vf < 0.0f) {
        dt = fmaxf(dtmin, dt * kFTDEC);
        damp = damp0;
        return mk3(0.f, 0.f, 0.f);
    }
    const float cv = 1.0f - damp;
    const float cf = damp * sqrtf(vv / (ff + kF2SAFE));
    dt = fminf(dtmax, dt * kFTINC);
    damp = damp * kFDAMP;
    return mk3(v.x * cv + f.x * cf, v.y * cv + f.y * cf, v.z * cv + f.z * cf);
}
// The integrator step before `v += f dt; pos += v dt`: update_FIRE, or the damped-velocity
// update `v *= (1 - damp)` of the reference's OPT_FIRE=0 build (kernels/AFM.cl:378-382), whose dt
// and damp stay at their initial values.
__device__ __forceinline__ f3 integrate(bool damped, f3 f, f3 v, float& dt, float& damp, float dtmin, float dtmax,
                                        float damp0) {
    if (damped) {
        const float cv = 1.0f - damp;
        return mk3(v.x * cv, v.y * cv, v.z * cv);
    }
    return fire(f, v, dt, damp, dtmin, dtmax, damp0);
}

__device__ __forceinline__ bool finite4(float4 v) {
    return isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w);
}

// lane -> scan point index; -1 when the lane has no point.
__device__ __forceinline__ int lane_point(int n_scan, int scan_ny) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (scan_ny > 0) {
        const int nxs = n_scan / scan_ny;
        const int tiles_y = (scan_ny + 7) >> 3;
        const int tx = wave / tiles_y, ty = wave - tx * tiles_y;
        const int ix = tx * 8 + (lane >> 3), iy = ty * 8 + (lane & 7);
        return (ix < nxs && iy < scan_ny) ? ix * scan_ny + iy : -1;
    }
    const int p = wave * 64 + lane;
    return p < n_scan ? p : -1;
}

// Scan point of this lane when G lanes share a point (point index, or -1), and the lane's
// index `sub` in its group.  Raster scans map a wave to a tile of (8/G) x 8 points.
template <int G>
__device__ __forceinline__ int group_point(int n_scan, int scan_ny, int& sub) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    sub = lane % G;
    const int q = lane / G;
    constexpr int TY = 8, TX = 8 / G;
    if (scan_ny > 0) {
        const int nxs = n_scan / scan_ny;
        const int tiles_y = (scan_ny + TY - 1) / TY;
        const int tx = wave / tiles_y, ty = wave - tx * tiles_y;
        const int ix = tx * TX + q / TY, iy = ty * TY + q % TY;
        return (ix < nxs && iy < scan_ny) ? ix * scan_ny + iy : -1;
    }
    const int p = wave * (64 / G) + q;
    return p < n_scan ? p : -1;
}

struct TiltArgs {
    int n_scan, nz, max_iter, check_before, scan_ny;
    int damped;             // 1: damped-velocity integrator (OPT_FIRE=0), 0: FIRE
    float f2conv;
    float A[3], B[3], C[3], dtip;
    float k[4], d0[4];
    float dt0, damp0, surf;
};

// Output staging of one lane: the newest (up to 4) states of its column are held in registers
// and stored as one run of back-to-back float4 stores when the run reaches a 64 B boundary of the
// output array or the column ends.  One float4 store per z-step at the column stride (the
// reference's (pixel, z) layout) left 64 B segments of L2 lines partly written for many
// microseconds; evicted in between, they were written back to memory several times over.
struct Stage4 {
    float4 b0, b1, b2, b3;   // oldest .. newest
    __device__ __forceinline__ void push(float4 v) { b0 = b1; b1 = b2; b2 = b3; b3 = v; }
    // stores the newest n (1..4) states ending at `last`
    __device__ __forceinline__ void flush(float4* last, int n) const {
        if (n >= 4) last[-3] = b0;
        if (n >= 3) last[-2] = b1;
        if (n >= 2) last[-1] = b2;
        last[0] = b3;
    }
};

// A linear operator along z applied to the lane's own Fz column after its relaxation (the df
// image): out[j] = sum_t W[j*taps + t] * Fz[clamp(start[j] + t)], accumulated with fmaf in t order
// (the accumulation of k_axis_operator for a single channel, so the results are those of the
// separate operator launch bit for bit).
struct DfOut {
    const float* W;
    const int* start;
    int taps, nout;
    bool lds;           // the launch holds each lane's Fz column in LDS
};

// relaxStrokesTilted (kernels/AFM.cl:478-562) for one scan point starting at tip anchor
// `start`; writes FEs[iz] (tip-rotated sample FE) and PPs[iz] (position, +/-iterations).
// `sample(x, y, z)` is the sample FE at a world point (trilinear grid sample, or the direct
// pair sum of relaxStrokesTiltedMorseDirect).  With the convergence test before the update
// the final resample is at the position just sampled, so its value is reused.
// G lanes share the point (same state in every lane): G = 1 stores FEs and PPs; otherwise lane
// sub 0 stores FEs and lane sub 1 PPs (the others only take part in the samples).  df (optional):
// the operator applied to the column after the relaxation (dfo, dfrow).  fz_lds (optional): this
// lane's Fz column in LDS (element iz at fz_lds[iz * 64]), kept by every lane of the group so that
// the G lanes split the df outputs; without it the FE-storing lane reads its stored column back.
// STAGE: outputs go through Stage4 runs (8 float4 registers for G = 1); false: one store per
// state, for kernels whose occupancy those registers would cost (the batched kernel shares its
// register budget with the grid tiles).
// LDSST (with STAGE, G = 1): the runs are staged in LDS instead of registers: `stage` is this lane's
// slot 0 of an 8 x 64 float4 area (slot k at stage[k * 64]; slots 0-3 FEs, 4-7 PPs), state k of a
// run in slot k (runs hold at most 4 states: they end at each 64 B boundary of the 16-byte aligned
// FEs).  Shifting registers through Stage4 made the compiler copy all 8 float4 at the loop's back
// edge on every FIRE step, and the store branch runs on nearly every step of a wave (some lane
// finishes a z-step on most of them).
// IDF: the tip frame is the identity (A, B, C = x, y, z).  The rotations then return their input
// and the z-only shifts (tip step, surface force) leave x and y alone, for finite values up to
// the sign of a zero, so they are skipped: the same results with ~30 fewer instructions per
// FIRE step.
template <int G, class SAMPLE, bool STAGE = true, bool IDF = false, bool LDSST = false>
__device__ __forceinline__ void relax_tilted_lane_src(SAMPLE& sample, f3 start, const TiltArgs& p,
                                                      float4* __restrict__ FEs, float4* __restrict__ PPs,
                                                      int sub = 0, const DfOut* dfo = nullptr,
                                                      float* __restrict__ dfrow = nullptr,
                                                      float* __restrict__ fz_lds = nullptr,
                                                      float4* __restrict__ stage = nullptr) {
    static_assert(!LDSST || (STAGE && G == 1), "LDS staging is built for one lane per point");
    const f3 A = ld3(p.A), B = ld3(p.B), C = ld3(p.C);
    const f3 dTip = mk3(C.x * p.dtip, C.y * p.dtip, C.z * p.dtip);
    const f3 d0r = IDF ? ld3(p.d0) : rotMatT(ld3(p.d0), A, B, C);
    f3 tip = start;
    f3 pos = mk3(tip.x + d0r.x, tip.y + d0r.y, tip.z + d0r.z);
    float dt = p.dt0, damp = p.damp0;
    const float dtmax = dt, dtmin = dtmax * 0.1f, damp0 = damp;
    const f3 fsurf = mk3(C.x * p.surf, C.y * p.surf, C.z * p.surf);
    f3 v = mk3(0.f, 0.f, 0.f);
    int it = 0, iz = 0;
    int o = 0;
    Stage4 s1, s2;       // G = 1: FEs, PPs; G > 1: the lane's own output
    int pend = 0;
    // The sample that gives a finished z-step's output (the resample at the final position) is
    // taken by the loop's own sample at the top of the next iteration (emit_next), so a wave whose
    // lanes finish their steps at different iterations runs one sample path, not a second one in
    // the branch of the finishing lanes; each lane's sequence of operations is unchanged.
    bool emit_next = false, conv_out = false;
    while (true) {
        const float4 fe = sample(pos.x, pos.y, pos.z);
        bool emit = emit_next;
        emit_next = false;
        if (!emit) {
            f3 f = mk3(fe.x, fe.y, fe.z);
            const f3 dp = mk3(pos.x - tip.x, pos.y - tip.y, pos.z - tip.z);
            if constexpr (IDF) {
                const f3 ft = tipForce(dp, p.k, p.d0);
                f = mk3(f.x + ft.x, f.y + ft.y, f.z + ft.z);
                f.z = f.z + p.surf;
            } else {
                const f3 ft = rotMatT(tipForce(rotMat(dp, A, B, C), p.k, p.d0), A, B, C);
                f = mk3(f.x + ft.x, f.y + ft.y, f.z + ft.z);
                f = mk3(f.x + fsurf.x, f.y + fsurf.y, f.z + fsurf.z);
            }
            const float ff = dot3(f, f);
            bool conv = false, done = false;
            if (p.check_before && ff < p.f2conv) {
                conv = done = true;
            } else {
                v = integrate(p.damped, f, v, dt, damp, dtmin, dtmax, damp0);
                v = mk3(v.x + f.x * dt, v.y + f.y * dt, v.z + f.z * dt);
                pos = mk3(pos.x + v.x * dt, pos.y + v.y * dt, pos.z + v.z * dt);
                if (!p.check_before && ff < p.f2conv) conv = done = true;
                else if (it + 1 >= p.max_iter) done = true;
            }
            ++it;
            if (done) {
                conv_out = conv;
                if (p.check_before && conv) emit = true;   // the output is this sample (no update)
                else emit_next = true;                     // the output is the sample at the new position
            }
        }
        if (emit) {
            const float4 fs = fe;
            const f3 fr = IDF ? mk3(fs.x, fs.y, fs.z) : rotMat(mk3(fs.x, fs.y, fs.z), A, B, C);
            const float4 feo = make_float4(fr.x, fr.y, fr.z, fs.w);
            const bool fin = finite4(fs) && isfinite(pos.x) && isfinite(pos.y) && isfinite(pos.z);
            const float4 ppo = make_float4(pos.x, pos.y, pos.z, (conv_out && fin) ? (float)it : -(float)it);
            if (fz_lds) fz_lds[iz * 64] = feo.z;
            if (!STAGE) {
                if (G == 1) {
                    FEs[o] = feo;
                    if (PPs) PPs[o] = ppo;
                } else if (sub == 0) {
                    FEs[o] = feo;
                } else if (sub == 1 && PPs) {
                    PPs[o] = ppo;
                }
            } else if constexpr (LDSST) {
                stage[pend * 64] = feo;
                stage[(4 + pend) * 64] = ppo;
            } else if (G == 1) {
                s1.push(feo);
                s2.push(ppo);
            } else {
                s1.push(sub == 1 ? ppo : feo);
            }
            ++pend;
            const bool last = iz + 1 >= p.nz;
            if (STAGE && LDSST && (last || (reinterpret_cast<uintptr_t>(FEs + o + 1) & 63) == 0)) {
                float4* dst = FEs + o - (pend - 1);
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (k < pend) dst[k] = stage[k * 64];
                if (PPs) {
                    float4* dq = PPs + o - (pend - 1);
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        if (k < pend) dq[k] = stage[(4 + k) * 64];
                }
                pend = 0;
            } else if (STAGE && !LDSST && (last || (reinterpret_cast<uintptr_t>(FEs + o + 1) & 63) == 0)) {
                if (G == 1) {
                    s1.flush(FEs + o, pend);
                    if (PPs) s2.flush(PPs + o, pend);
                } else if (sub == 0) {
                    s1.flush(FEs + o, pend);
                } else if (sub == 1 && PPs) {
                    s1.flush(PPs + o, pend);
                }
                pend = 0;
            }
            ++o;
            if constexpr (IDF) {
                tip.z = tip.z + p.dtip;
                pos.z = pos.z + p.dtip;
            } else {
                tip = mk3(tip.x + dTip.x, tip.y + dTip.y, tip.z + dTip.z);
                pos = mk3(pos.x + dTip.x, pos.y + dTip.y, pos.z + dTip.z);
            }
            v = mk3(0.f, 0.f, 0.f);
            it = 0;
            if (++iz >= p.nz) break;
        }
    }
    if (dfo && fz_lds) {         // the lane's own copy of the column; the group splits the outputs
        for (int j = sub; j < dfo->nout; j += G) {
            const int s0 = dfo->start[j];
            float acc = 0.f;
            for (int t = 0; t < dfo->taps; ++t) {
                const int m = min(max(s0 + t, 0), p.nz - 1);
                acc = fmaf(dfo->W[j * dfo->taps + t], fz_lds[m * 64], acc);
            }
            dfrow[j] = acc;
        }
    } else if (dfo && sub == 0) {  // the column just stored by this lane
        for (int j = 0; j < dfo->nout; ++j) {
            const int s0 = dfo->start[j];
            float acc = 0.f;
            for (int t = 0; t < dfo->taps; ++t) {
                const int m = min(max(s0 + t, 0), p.nz - 1);
                acc = fmaf(dfo->W[j * dfo->taps + t], FEs[m].z, acc);
            }
            dfrow[j] = acc;
        }
    }
}

// The tip frame is the identity, the grid's inverse cell has no off-diagonal terms and its stored
// planes are a small_grid: the relaxation may take its FAST form (same results).
__host__ __device__ inline bool fast_frame(const TiltArgs& p, const GridDev& g) {
    return p.A[0] == 1.f && p.A[1] == 0.f && p.A[2] == 0.f && p.B[0] == 0.f && p.B[1] == 1.f && p.B[2] == 0.f &&
           p.C[0] == 0.f && p.C[1] == 0.f && p.C[2] == 1.f && g.da[1] == 0.f && g.da[2] == 0.f && g.db[0] == 0.f &&
           g.db[2] == 0.f && g.dc[0] == 0.f && g.dc[1] == 0.f && small_grid(g);
}

template <bool STAGE = true, bool FAST = false, bool LDSST = false>
__device__ __forceinline__ void relax_tilted_lane(const GridDev& g, f3 start, const TiltArgs& p,
                                                  float4* __restrict__ FEs, float4* __restrict__ PPs,
                                                  const DfOut* dfo = nullptr, float* __restrict__ dfrow = nullptr,
                                                  float* __restrict__ fz_lds = nullptr,
                                                  float4* __restrict__ stage = nullptr) {
    Cell cell;
    cell.reset();
    auto smp = [&](float x, float y, float z) { return sample<FAST>(g, cell, x, y, z); };
    relax_tilted_lane_src<1, decltype(smp), STAGE, FAST, LDSST>(smp, start, p, FEs, PPs, 0, dfo, dfrow, fz_lds,
                                                                stage);
}

}  // namespace rc
}  // namespace ppm
