// ppm_relax.hip — probe relaxation along z strokes for gfx950 (stage b).
//
// Built with -ffp-contract=off: every FP32 operation is evaluated exactly as written,
// in the order of the reference kernels (no fused multiply-add), so results are
// reproducible against any IEEE-754 FP32 evaluation of the same sequence.
//
// Semantics (one lane = one scan point, z descent with warm start):
//   relaxStrokesTilted    kernels/AFM.cl:478-562
//   relaxStrokes          kernels/AFM.cl:345-394
//   getFEinStrokesTilted  kernels/AFM.cl:210-233
//   interpFE + sampler_1  kernels/AFM.cl:49, :100-105 — normalized coordinates,
//                         REPEAT addressing on all three axes, linear filter; the
//                         trilinear weights are exact FP32 (the OpenCL 1.2 linear-filter
//                         formula), combined in the nested order of
//                         read_imagef_trilin_ (kernels/AFM.cl:81-97)
//   tipForce              kernels/AFM.cl:53-58
//   update_FIRE           kernels/AFM.cl:121-157
//
// Design:
//   * each lane keeps the 8 float4 corners of the grid cell it last sampled in
//     registers and reloads them only when the probe moves into another cell: near
//     convergence a FIRE iteration touches no memory at all;
//   * lanes advance through their z slices independently (one merged loop over FIRE
//     iterations), so a wavefront is not held at each slice by its slowest lane;
//   * optional 8x8 pixel tiles per wavefront (raster scans) keep the cells of
//     neighbouring lanes in the same cache lines.
#include "ppm_relax_core.hpp"

namespace ppm {
namespace {

using namespace rc;

// State gates of the adjoint, the values of kernels/AFM.cl:1778, :1826, :1836.
constexpr double kResidualGate = 5e-4, kCondGate = 1e6, kStableGate = 1e-5;

// LDS of the staged output runs of one workgroup (one wave): 8 float4 slots per lane.
constexpr size_t kStageBytes = 8 * 64 * sizeof(float4);

// relaxStrokesTilted with PP telemetry; G lanes per scan point (G = 1, 2, 4): with G > 1 each
// lane interpolates 4/G channels and the group shares the FIRE logic, results identical to G = 1.
// df (optional, with dfo): the df operator applied to each point's Fz column, (n_scan, dfo.nout).
// FAST: identity tip frame and grid axes along the world axes (the form of every untilted scan),
// which skips the rotations and the off-axis texture-coordinate terms (same results).
// LDSST (G = 1): the FEs and PPs runs are staged in LDS (kStageBytes per workgroup) instead of
// registers: 12.7 % fewer VALU instructions on the flake scan, 91 instead of 114 VGPRs.
template <int G, bool FAST, bool LDSST = false>
__global__ __launch_bounds__(64) void k_relax_tilted(GridDev g, const float4* __restrict__ pts, TiltArgs p,
                                                     float4* __restrict__ FEs, float4* __restrict__ PPs,
                                                     DfOut dfo, float* __restrict__ df) {
    static_assert(!LDSST || G == 1, "LDS staging is built for one lane per point");
    int sub;
    const int ip = group_point<G>(p.n_scan, p.scan_ny, sub);
    if (ip < 0) return;                       // whole groups leave together
    const float4 P = pts[ip];
    const size_t o = (size_t)ip * p.nz;
    const DfOut* dp = df ? &dfo : nullptr;
    float* row = df ? df + (size_t)ip * dfo.nout : nullptr;
    // LDS: with LDSST the output runs of the wave (8 x 64 float4: FEs slots 0-3, PPs 4-7), then, when
    // the launch gives it room, the Fz columns (nz x 64 floats)
    extern __shared__ float4 s_dyn[];
    float4* stage = LDSST ? s_dyn + (threadIdx.x & 63) : nullptr;
    float* s_fz = reinterpret_cast<float*>(s_dyn + (LDSST ? kStageBytes / sizeof(float4) : 0));
    float* col = (df && dfo.lds) ? s_fz + (threadIdx.x & 63) : nullptr;
    if constexpr (G == 1) {
        relax_tilted_lane<true, FAST, LDSST>(g, mk3(P.x, P.y, P.z), p, FEs + o, PPs ? PPs + o : nullptr, dp, row,
                                             col, stage);
    } else {
        GroupCell<G> cell;
        cell.reset();
        auto smp = [&](float x, float y, float z) { return sample_group<G, FAST>(g, cell, sub, x, y, z); };
        relax_tilted_lane_src<G, decltype(smp), true, FAST>(smp, mk3(P.x, P.y, P.z), p, FEs + o,
                                                            PPs ? PPs + o : nullptr, sub, dp, row, col);
    }
}


// Lanes per scan point of k_relax_tilted / k_relax_strokes: the largest G <= 4 that keeps the
// launch within 2.5 waves per SIMD (10 per compute unit).  A group replicates the per-point FIRE
// and spring arithmetic in its G lanes, so past that occupancy the extra waves cost VALU issue
// slots instead of hiding latency.  Measured on the flake grid (G = 4, 2, 1; bit-identical
// results): 16,384 pixels 1.92, 2.31, 2.64 ms; 40,000 pixels (2500 waves at G = 4) 2.16, 2.35,
// 2.76; 43,264 (2704) 2.72, 2.40, 2.83; 65,536 3.72, 2.46, 2.73; 73,984 (2312 waves at G = 2)
// 4.06, 2.53, 2.70; 82,944 (2592) 4.39, 2.89, 2.71.
int relax_group(int n_scan) {
    const long long max_lanes = 10LL * device_cus() * 64;
    int g = 1;
    while (g < 4 && (long long)n_scan * g * 2 <= max_lanes) g *= 2;
    return g;
}

int group_waves(int n_scan, int scan_ny, int g) {
    if (scan_ny > 0) {
        const int nxs = n_scan / scan_ny, tx = 8 / g;
        return ((nxs + tx - 1) / tx) * ((scan_ny + 7) / 8);
    }
    return (int)(((long long)n_scan * g + 63) / 64);
}

struct StrokeArgs {
    int n_scan, nz, max_iter, scan_ny;
    int damped;             // 1: damped-velocity integrator (OPT_FIRE=0), 0: FIRE
    float f2conv;
    float dTip[3];
    float k[4], d0[4];
    float dt0, damp0;
};

// relaxStrokes with optional displacement and PP outputs, for one scan point; `smp(x,y,z)` is the
// trilinear sample (one lane's cell cache, or a lane group's channel-split sample), `writer` the
// lane that stores (all lanes of a group hold the same state).
template <class SAMPLE>
__device__ __forceinline__ void strokes_lane(SAMPLE& smp, int ip, const float4* __restrict__ pts,
                                             const StrokeArgs& p, float4* __restrict__ FEs,
                                             float4* __restrict__ disps, float4* __restrict__ PPs, bool writer) {
    const f3 dTip = ld3(p.dTip);
    const float4 P = pts[ip];
    f3 tip = mk3(P.x, P.y, P.z);
    f3 pos = mk3(tip.x + p.d0[0], tip.y + p.d0[1], tip.z + p.d0[2]);
    float dt = p.dt0, damp = p.damp0;
    const float dtmax = dt, dtmin = dtmax * 0.1f, damp0 = damp;
    f3 v = mk3(0.f, 0.f, 0.f);
    int it = 0, iz = 0;
    size_t o = (size_t)ip * p.nz;
    while (true) {
        const float4 fe = smp(pos.x, pos.y, pos.z);
        const f3 ft = tipForce(mk3(pos.x - tip.x, pos.y - tip.y, pos.z - tip.z), p.k, p.d0);
        const f3 f = mk3(fe.x + ft.x, fe.y + ft.y, fe.z + ft.z);
        const float ff = dot3(f, f);
        v = integrate(p.damped, f, v, dt, damp, dtmin, dtmax, damp0);
        v = mk3(v.x + f.x * dt, v.y + f.y * dt, v.z + f.z * dt);
        pos = mk3(pos.x + v.x * dt, pos.y + v.y * dt, pos.z + v.z * dt);
        const bool conv = ff < p.f2conv;
        ++it;
        if (conv || it >= p.max_iter) {
            if (writer) {
                FEs[o] = fe;
                if (disps)
                    disps[o] = make_float4(pos.x - (tip.x + p.d0[0]), pos.y - (tip.y + p.d0[1]),
                                           pos.z - (tip.z + p.d0[2]), 0.f);
                if (PPs) {
                    const bool fin = finite4(fe) && isfinite(pos.x) && isfinite(pos.y) && isfinite(pos.z);
                    PPs[o] = make_float4(pos.x, pos.y, pos.z, (conv && fin) ? (float)it : -(float)it);
                }
            }
            ++o;
            tip = mk3(tip.x + dTip.x, tip.y + dTip.y, tip.z + dTip.z);
            pos = mk3(pos.x + dTip.x, pos.y + dTip.y, pos.z + dTip.z);
            v = mk3(0.f, 0.f, 0.f);
            it = 0;
            if (++iz >= p.nz) break;
        }
    }
}

// relaxStrokes; G lanes per scan point as k_relax_tilted (results identical to G = 1).
template <int G>
__global__ __launch_bounds__(64) void k_relax_strokes(GridDev g, const float4* __restrict__ pts, StrokeArgs p,
                                                      float4* __restrict__ FEs, float4* __restrict__ disps,
                                                      float4* __restrict__ PPs) {
    int sub;
    const int ip = group_point<G>(p.n_scan, p.scan_ny, sub);
    if (ip < 0) return;                       // whole groups leave together
    if constexpr (G == 1) {
        Cell cell;
        cell.reset();
        auto smp = [&](float x, float y, float z) { return sample(g, cell, x, y, z); };
        strokes_lane(smp, ip, pts, p, FEs, disps, PPs, true);
    } else {
        GroupCell<G> cell;
        cell.reset();
        auto smp = [&](float x, float y, float z) { return sample_group<G>(g, cell, sub, x, y, z); };
        strokes_lane(smp, ip, pts, p, FEs, disps, PPs, sub == 0);
    }
}

struct Strokes2DArgs {
    int n_scan, nz, max_iter, scan_ny;
    float f2conv, k_lat, dh, dt, damp;
};

// relaxStrokes2D (kernels/AFM.cl:396-450) for one scan point: the probe height is fixed per slice
// (pz = az - iz*dh), only x,y relax, from the anchor at every slice, by damped velocity steps
// against the lateral spring -k_lat (p - anchor); the |F_xy|^2 test follows the update.  FEs: the
// sample at the final position; disps: (x - ax, y - ay, 0, 0); PPs (telemetry): position and
// +iterations (converged) / -iterations.
template <class SAMPLE>
__device__ __forceinline__ void strokes2d_lane(SAMPLE& smp, int ip, const float4* __restrict__ pts,
                                               const Strokes2DArgs& p, float4* __restrict__ FEs,
                                               float4* __restrict__ disps, float4* __restrict__ PPs, bool writer) {
    const float4 P = pts[ip];
    const float ax = P.x, ay = P.y, az = P.z;
    const float cv = 1.0f - p.damp;
    size_t o = (size_t)ip * p.nz;
    for (int iz = 0; iz < p.nz; ++iz, ++o) {
        const float pz = az - (float)iz * p.dh;
        float px = ax, py = ay, vx = 0.f, vy = 0.f;
        int it = 0;
        bool conv = false;
        for (it = 0; it < p.max_iter; ++it) {
            const float4 fe = smp(px, py, pz);
            const float fx = fe.x - p.k_lat * (px - ax);
            const float fy = fe.y - p.k_lat * (py - ay);
            vx = vx * cv + fx * p.dt;
            vy = vy * cv + fy * p.dt;
            px += vx * p.dt;
            py += vy * p.dt;
            if ((fx * fx + fy * fy) < p.f2conv) { conv = true; break; }
        }
        const int iters = conv ? it + 1 : p.max_iter;
        const float4 fe = smp(px, py, pz);
        if (writer) {
            FEs[o] = fe;
            if (disps) disps[o] = make_float4(px - ax, py - ay, 0.f, 0.f);
            if (PPs) {
                const bool fin = finite4(fe) && isfinite(px) && isfinite(py);
                PPs[o] = make_float4(px, py, pz, (conv && fin) ? (float)iters : -(float)iters);
            }
        }
    }
}

// relaxStrokes2D; G lanes per scan point as k_relax_tilted (results identical to G = 1).
template <int G>
__global__ __launch_bounds__(64) void k_relax_strokes_2d(GridDev g, const float4* __restrict__ pts, Strokes2DArgs p,
                                                         float4* __restrict__ FEs, float4* __restrict__ disps,
                                                         float4* __restrict__ PPs) {
    int sub;
    const int ip = group_point<G>(p.n_scan, p.scan_ny, sub);
    if (ip < 0) return;                       // whole groups leave together
    if constexpr (G == 1) {
        Cell cell;
        cell.reset();
        auto smp = [&](float x, float y, float z) { return sample(g, cell, x, y, z); };
        strokes2d_lane(smp, ip, pts, p, FEs, disps, PPs, true);
    } else {
        GroupCell<G> cell;
        cell.reset();
        auto smp = [&](float x, float y, float z) { return sample_group<G>(g, cell, sub, x, y, z); };
        strokes2d_lane(smp, ip, pts, p, FEs, disps, PPs, sub == 0);
    }
}

struct RawArgs {
    int n_scan, nz, scan_ny;
    float A[3], B[3], C[3];
    float dTip[3], d0[3];
};

// getFEinStrokesTilted: unrelaxed samples at point + dpos0 + iz*dTip, rotated into the tip frame.
__global__ __launch_bounds__(64) void k_fe_in_strokes(GridDev g, const float4* __restrict__ pts, RawArgs p,
                                                      float4* __restrict__ FEs) {
    const int ip = lane_point(p.n_scan, p.scan_ny);
    if (ip < 0) return;
    const f3 A = ld3(p.A), B = ld3(p.B), C = ld3(p.C);
    const float4 P = pts[ip];
    f3 pos = mk3(P.x + p.d0[0], P.y + p.d0[1], P.z + p.d0[2]);
    Cell cell;
    cell.reset();
    for (int iz = 0; iz < p.nz; ++iz) {
        const float4 fe = sample(g, cell, pos.x, pos.y, pos.z);
        const f3 fr = rotMat(mk3(fe.x, fe.y, fe.z), A, B, C);
        FEs[(size_t)ip * p.nz + iz] = make_float4(fr.x, fr.y, fr.z, fe.w);
        pos = mk3(pos.x + p.dTip[0], pos.y + p.dTip[1], pos.z + p.dTip[2]);
    }
}

// ── relaxed-state calculus: total force, its Jacobian, Newton polish, IFT adjoint ──────
//
// At a probe position q the total force is G(q) = F(q) + R^T g(R(q - tip)) + C*surf, with F the
// trilinear sample of the grid, g = tipForce in the tip frame R = rows (A,B,C).  Inside the grid
// cell that contains q, F is trilinear, so J = dG/dq = J_F + R^T (dg/dd) R is available in
// closed form (J_F from the 8 corners, dg/dd = diag(k) + k_w(1 - L/r) I + (k_w L / r^3) d d^T).
struct StateArgs {
    int n_scan, nz, steps;
    float A[3], B[3], C[3], dtip;
    float k[4], d0[4];
    float surf;
};

struct StateEval {
    size_t idx[8];
    float w[8];
    float4 fe;          // trilinear sample at q, rounded to FP32
    double JF[3][3];    // dF_xyz/dq inside the cell
    double gE[3];       // dE/dq inside the cell
    double G[3];        // total force
    double J[3][3];     // dG/dq
};

__device__ __forceinline__ f3 tip_anchor(const float4* pts, const StateArgs& p, int ip, int iz) {
    const f3 C = ld3(p.C);
    const f3 dTip = mk3(C.x * p.dtip, C.y * p.dtip, C.z * p.dtip);
    const float4 P0 = pts[ip];
    f3 tip = mk3(P0.x, P0.y, P0.z);
    for (int k = 0; k < iz; ++k) tip = mk3(tip.x + dTip.x, tip.y + dTip.y, tip.z + dTip.z);   // as the forward
    return tip;
}

// REPEAT addressing of one axis in FP64: the cell (i0, i1) and the weight a of the upper corner,
// the texel-centre convention of axis_frac / axis_index without their FP32 rounding of the coordinate.
__device__ __forceinline__ void axis_repeat_d(double s, int n, int& i0, int& i1, double& a) {
    const double u = (s - floor(s)) * (double)n;
    const double t = u - 0.5;
    double fl = floor(t);
    a = t - fl;
    fl = fmin(fmax(fl, -1.0), (double)n);      // keeps non-finite coordinates in range
    int i = (int)fl;
    int j = i + 1;
    if (i < 0) i += n;
    if (j > n - 1) j -= n;
    i0 = min(max(i, 0), n - 1);
    i1 = min(max(j, 0), n - 1);
}

// The trilinear model of the FP32 grid at an FP64 position q: sampling coordinates, weights and
// interpolation in FP64 (the corner values are the grid's FP32 values), so the model is the
// continuous piecewise-trilinear field whose equilibria the Newton polish solves for and whose
// derivatives the adjoint uses (an FP32 sample would quantize q to ~1e-6 A through its texture
// coordinates).
__device__ void eval_state(const GridDev& g, const StateArgs& p, f3 tip, double qx, double qy, double qz,
                           StateEval& e) {
    const double sa = qx * (double)g.da[0] + qy * (double)g.da[1] + qz * (double)g.da[2] + (double)g.da[3];
    const double sb = qx * (double)g.db[0] + qy * (double)g.db[1] + qz * (double)g.db[2] + (double)g.db[3];
    const double sc = qx * (double)g.dc[0] + qy * (double)g.dc[1] + qz * (double)g.dc[2] + (double)g.dc[3];
    int i0, i1, j0, j1, k0, k1;
    double a, b, c;
    axis_repeat_d(sa, g.nx, i0, i1, a);
    axis_repeat_d(sb, g.ny, j0, j1, b);
    axis_repeat_d(sc, g.nz, k0, k1, c);
    const size_t sx = (size_t)g.nx, sxy = (size_t)g.nx * g.ny;
    e.idx[0] = (size_t)k0 * sxy + (size_t)j0 * sx + i0; e.idx[1] = (size_t)k0 * sxy + (size_t)j0 * sx + i1;
    e.idx[2] = (size_t)k0 * sxy + (size_t)j1 * sx + i0; e.idx[3] = (size_t)k0 * sxy + (size_t)j1 * sx + i1;
    e.idx[4] = (size_t)k1 * sxy + (size_t)j0 * sx + i0; e.idx[5] = (size_t)k1 * sxy + (size_t)j0 * sx + i1;
    e.idx[6] = (size_t)k1 * sxy + (size_t)j1 * sx + i0; e.idx[7] = (size_t)k1 * sxy + (size_t)j1 * sx + i1;
    Cell C;
#pragma unroll
    for (int m = 0; m < 8; ++m) C.c[m] = g.data[e.idx[m]];
    const double ma = 1.0 - a, mb = 1.0 - b, mc = 1.0 - c;
    const double w[8] = {ma * mb * mc, a * mb * mc, ma * b * mc, a * b * mc,
                         ma * mb * c,  a * mb * c,  ma * b * c,  a * b * c};
#pragma unroll
    for (int m = 0; m < 8; ++m) e.w[m] = (float)w[m];
    double fd[4], dFa[4], dFb[4], dFc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double f[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) f[m] = (double)(&C.c[m].x)[q];
        fd[q] = ((f[0] * ma + f[1] * a) * mb + (f[2] * ma + f[3] * a) * b) * mc +
                ((f[4] * ma + f[5] * a) * mb + (f[6] * ma + f[7] * a) * b) * c;
        dFa[q] = ((f[1] - f[0]) * mb + (f[3] - f[2]) * b) * mc + ((f[5] - f[4]) * mb + (f[7] - f[6]) * b) * c;
        dFb[q] = ((f[2] - f[0]) * ma + (f[3] - f[1]) * a) * mc + ((f[6] - f[4]) * ma + (f[7] - f[5]) * a) * c;
        dFc[q] = ((f[4] - f[0]) * ma + (f[5] - f[1]) * a) * mb + ((f[6] - f[2]) * ma + (f[7] - f[3]) * a) * b;
    }
    e.fe = make_float4((float)fd[0], (float)fd[1], (float)fd[2], (float)fd[3]);
    double ta[3], tb[3], tc[3];
    for (int j = 0; j < 3; ++j) {
        ta[j] = (double)g.nx * (double)g.da[j];
        tb[j] = (double)g.ny * (double)g.db[j];
        tc[j] = (double)g.nz * (double)g.dc[j];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) e.JF[i][j] = dFa[i] * ta[j] + dFb[i] * tb[j] + dFc[i] * tc[j];
    for (int j = 0; j < 3; ++j) e.gE[j] = dFa[3] * ta[j] + dFb[3] * tb[j] + dFc[3] * tc[j];
    const double R[3][3] = {{p.A[0], p.A[1], p.A[2]}, {p.B[0], p.B[1], p.B[2]}, {p.C[0], p.C[1], p.C[2]}};
    const double dw[3] = {qx - (double)tip.x, qy - (double)tip.y, qz - (double)tip.z};
    double d[3];
    for (int i = 0; i < 3; ++i) d[i] = R[i][0] * dw[0] + R[i][1] * dw[1] + R[i][2] * dw[2];
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    r = r > 1e-10 ? r : 1e-10;
    const double phi = (double)p.k[3] * (1.0 - (double)p.d0[3] / r);
    const double beta = (double)p.k[3] * (double)p.d0[3] / (r * r * r);
    double gt[3], Jt[3][3];
    for (int i = 0; i < 3; ++i) {
        gt[i] = (d[i] - (double)p.d0[i]) * (double)p.k[i] + d[i] * phi;
        for (int j = 0; j < 3; ++j) Jt[i][j] = beta * d[i] * d[j] + (i == j ? (double)p.k[i] + phi : 0.0);
    }
    const double fs[3] = {fd[0], fd[1], fd[2]};
    for (int i = 0; i < 3; ++i) {
        e.G[i] = fs[i] + R[0][i] * gt[0] + R[1][i] * gt[1] + R[2][i] * gt[2] + (double)p.C[i] * (double)p.surf;
        for (int j = 0; j < 3; ++j) {
            double acc = e.JF[i][j];
            for (int m = 0; m < 3; ++m)
                for (int n = 0; n < 3; ++n) acc += R[m][i] * Jt[m][n] * R[n][j];
            e.J[i][j] = acc;
        }
    }
}

// cof[i][j] = cofactor of J; returns det(J) and whether J is numerically regular
__device__ __forceinline__ bool cofactors(const double J[3][3], double cof[3][3], double& det) {
    cof[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    cof[0][1] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    cof[0][2] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    cof[1][0] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    cof[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    cof[1][2] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
    cof[2][0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    cof[2][1] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    cof[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    det = J[0][0] * cof[0][0] + J[0][1] * cof[0][1] + J[0][2] * cof[0][2];
    double scale = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) scale = fmax(scale, fabs(J[i][j]));
    return scale > 0.0 && fabs(det) > 1e-12 * scale * scale * scale;
}

__device__ __forceinline__ double norm3(const double v[3]) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// Newton polish of converged probe positions (FP64 steps on the trilinear model): a step
// q <- q - J^{-1} G is kept only if it lowers |G|.  FEs are resampled at the final position.
__global__ __launch_bounds__(256) void k_relax_polish(GridDev g, const float4* __restrict__ pts, StateArgs p,
                                                      float4* __restrict__ PPs, float4* __restrict__ FEs) {
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (size_t)p.n_scan * p.nz) return;
    const float4 PP = PPs[s];
    if (!(PP.w > 0.f)) return;
    const int ip = (int)(s / p.nz);
    const int iz = (int)(s - (size_t)ip * p.nz);
    const f3 tip = tip_anchor(pts, p, ip, iz);
    double q[3] = {PP.x, PP.y, PP.z};
    StateEval e;
    eval_state(g, p, tip, q[0], q[1], q[2], e);
    double gn = norm3(e.G);
    for (int it = 0; it < p.steps; ++it) {
        double cof[3][3], det;
        if (!cofactors(e.J, cof, det)) break;
        double dq[3];
        for (int i = 0; i < 3; ++i) dq[i] = -(cof[0][i] * e.G[0] + cof[1][i] * e.G[1] + cof[2][i] * e.G[2]) / det;
        const double qt[3] = {q[0] + dq[0], q[1] + dq[1], q[2] + dq[2]};
        StateEval et;
        eval_state(g, p, tip, qt[0], qt[1], qt[2], et);
        const double gt = norm3(et.G);
        if (!(gt < gn)) break;
        q[0] = qt[0]; q[1] = qt[1]; q[2] = qt[2];
        e = et;
        gn = gt;
    }
    PPs[s] = make_float4((float)q[0], (float)q[1], (float)q[2], PP.w);
    if (FEs) {
        const f3 fr = rotMat(mk3(e.fe.x, e.fe.y, e.fe.z), ld3(p.A), ld3(p.B), ld3(p.C));
        FEs[s] = make_float4(fr.x, fr.y, fr.z, e.fe.w);
    }
}

// Smallest eigenvalue of a symmetric 3x3 matrix: trigonometric closed form of
// O. K. Smith, "Eigenvalues of a symmetric 3x3 matrix", Commun. ACM 4(4):168 (1961).
__device__ __forceinline__ double eigmin_sym3(const double S[3][3]) {
    const double p1 = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
    if (p1 == 0.0) return fmin(S[0][0], fmin(S[1][1], S[2][2]));
    const double q = (S[0][0] + S[1][1] + S[2][2]) / 3.0;
    const double a = S[0][0] - q, b = S[1][1] - q, c = S[2][2] - q;
    const double p = sqrt((a * a + b * b + c * c + 2.0 * p1) / 6.0);
    const double ip = 1.0 / p;
    const double B00 = a * ip, B11 = b * ip, B22 = c * ip;
    const double B01 = S[0][1] * ip, B02 = S[0][2] * ip, B12 = S[1][2] * ip;
    const double detB = B00 * (B11 * B22 - B12 * B12) - B01 * (B01 * B22 - B12 * B02) + B02 * (B01 * B12 - B11 * B02);
    const double r = fmin(fmax(0.5 * detB, -1.0), 1.0);
    const double phi = acos(r) / 3.0;
    return q + 2.0 * p * cos(phi + 2.0943951023931957);      // + 2 pi / 3: the smallest root
}

// Infinity-norm condition number ||J|| ||J^{-1}|| with J^{-1} = cof^T / det.
__device__ __forceinline__ double cond_inf(const double J[3][3], const double cof[3][3], double det) {
    double jn = 0.0, an = 0.0;
    for (int i = 0; i < 3; ++i) {
        jn = fmax(jn, fabs(J[i][0]) + fabs(J[i][1]) + fabs(J[i][2]));
        an = fmax(an, fabs(cof[0][i]) + fabs(cof[1][i]) + fabs(cof[2][i]));
    }
    return jn * an / fmax(fabs(det), 1e-300);
}

// Implicit-function adjoint of one relaxed state.  The output O = (R F(q*), E(q*)); for an
// upstream cotangent u = dL/dO:
//   b      = (dO/dq)^T u = J_F^T R^T u_xyz + u_w grad E
//   J^T l  = b                                  (3x3, FP64)
//   dL/dG_c.xyz += w_c (R^T u_xyz - l),   dL/dG_c.w += w_c u_w     for the 8 corners c
// (dL/dG = u . dO/dG - l^T dG/dG).  A state is used only if it passes the gates of the
// reference's direct-path adjoint (kernels/AFM.cl:1778, :1811-1840): residual |G(q*)| < 5e-4,
// finite, regular J with condition number < 1e6, and a stable minimum (smallest eigenvalue of
// the symmetric part of -J above 1e-5).  Status bits (0 = used): 1 residual, 2 non-finite,
// 4 singular or ill-conditioned, 8 unstable.  A gated state contributes nothing.
__device__ __forceinline__ void atomic_add4(float* p, float4 v) {
    atomicAdd(p + 0, v.x);
    atomicAdd(p + 1, v.y);
    atomicAdd(p + 2, v.z);
    atomicAdd(p + 3, v.w);
}

// Consecutive states of a scan point sit in consecutive lanes; when the probe descends by one grid
// plane per step, the lower corner plane of a state is the upper corner plane of the next one.
// The next lane then adds those 4 corner values to its own before it scatters, and this lane
// skips them: half the atomics of such a scan (the atomics, executed at the memory side so that
// the 8 XCDs see one sum, bound the kernel).
__global__ __launch_bounds__(256) void k_relax_adjoint(GridDev g, const float4* __restrict__ pts, StateArgs p,
                                                       const float4* __restrict__ PPs, const float4* __restrict__ U,
                                                       float* __restrict__ gcot, float4* __restrict__ lam) {
    const size_t n = (size_t)p.n_scan * p.nz;
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool inrange = s < n;        // every lane stays to the end: the corner exchange is a wave shuffle
    int status = 1;
    StateEval e;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        e.idx[m] = 0;
        e.w[m] = 0.f;
    }
    float4 u = make_float4(0.f, 0.f, 0.f, 0.f);
    float gx = 0.f, gy = 0.f, gz = 0.f;
    if (inrange) {
        const int ip = (int)(s / p.nz);
        const int iz = (int)(s - (size_t)ip * p.nz);
        const f3 tip = tip_anchor(pts, p, ip, iz);
        const float4 PP = PPs[s];
        u = U[s];
        status = 0;
        eval_state(g, p, tip, PP.x, PP.y, PP.z, e);
        if (!(PP.w > 0.f) || !(norm3(e.G) < kResidualGate)) status |= 1;
        const double R[3][3] = {{p.A[0], p.A[1], p.A[2]}, {p.B[0], p.B[1], p.B[2]}, {p.C[0], p.C[1], p.C[2]}};
        double ru[3];
        for (int j = 0; j < 3; ++j) ru[j] = R[0][j] * u.x + R[1][j] * u.y + R[2][j] * u.z;
        double bv[3];
        for (int j = 0; j < 3; ++j)
            bv[j] = e.JF[0][j] * ru[0] + e.JF[1][j] * ru[1] + e.JF[2][j] * ru[2] + (double)u.w * e.gE[j];
        double cof[3][3], det;
        double l[3] = {0.0, 0.0, 0.0};
        if (!cofactors(e.J, cof, det) || !(cond_inf(e.J, cof, det) < kCondGate)) {
            status |= 4;
        } else {
            // J^T l = b:  (J^T)^{-1} = adj(J^T)/det = cof(J)/det
            for (int i = 0; i < 3; ++i) l[i] = (cof[i][0] * bv[0] + cof[i][1] * bv[1] + cof[i][2] * bv[2]) / det;
        }
        double S[3][3];                      // symmetric part of the Hessian -J
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) S[i][j] = -0.5 * (e.J[i][j] + e.J[j][i]);
        if (!(eigmin_sym3(S) > kStableGate)) status |= 8;
        if (!(isfinite(l[0]) && isfinite(l[1]) && isfinite(l[2]) && isfinite(PP.x) && isfinite(PP.y) &&
              isfinite(PP.z) && isfinite(bv[0]) && isfinite(bv[1]) && isfinite(bv[2])))
            status |= 2;
        if (status) l[0] = l[1] = l[2] = 0.0;
        lam[s] = make_float4((float)l[0], (float)l[1], (float)l[2], (float)status);
        gx = (float)(ru[0] - l[0]);
        gy = (float)(ru[1] - l[1]);
        gz = (float)(ru[2] - l[2]);
    }
    const bool give = status == 0;           // this lane scatters its corners
    float4 c[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const float w = give ? e.w[m] : 0.f;
        c[m] = make_float4(w * gx, w * gy, w * gz, w * u.w);
    }
    unsigned long long key[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) key[m] = give ? (unsigned long long)e.idx[m] : ~0ull;
    // lower plane of the previous lane (lane - 1 of the same wave); every lane executes every
    // shuffle (a shuffle under a divergent condition would read lanes that did not take part)
    const int lane = threadIdx.x & 63;
    const int prev_give = __shfl_up((int)give, 1, 64);
    bool take = give && lane > 0 && prev_give != 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const unsigned long long prev = __shfl_up(key[m], 1, 64);
        take = take && prev == (unsigned long long)e.idx[4 + m];
    }
    float4 prevc[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        prevc[m].x = __shfl_up(c[m].x, 1, 64);
        prevc[m].y = __shfl_up(c[m].y, 1, 64);
        prevc[m].z = __shfl_up(c[m].z, 1, 64);
        prevc[m].w = __shfl_up(c[m].w, 1, 64);
    }
    const int next_take = __shfl_down((int)take, 1, 64);
    const bool next_took = lane < 63 && next_take != 0;
    if (!give) return;
    if (take) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            c[4 + m].x += prevc[m].x;
            c[4 + m].y += prevc[m].y;
            c[4 + m].z += prevc[m].z;
            c[4 + m].w += prevc[m].w;
        }
    }
#pragma unroll
    for (int m = 0; m < 8; ++m)
        if (m >= 4 || !next_took) atomic_add4(gcot + 4 * e.idx[m], c[m]);
}

// Per-state diagnostics of relaxed states without a cotangent: the adjoint_diag record of the
// reference's direct-path state adjoint (kernels/AFM.cl:1774-1844) for the trilinear grid model:
// (|G(q*)|, smallest eigenvalue of the symmetric part of -J, infinity-norm condition number of J,
// status bits of k_relax_adjoint).  A caller uses it to find states the adjoint would gate, or
// probes close to a fold of their equilibrium branch (small smallest eigenvalue), before it
// forms a loss.
__global__ __launch_bounds__(256) void k_relax_state_diag(GridDev g, const float4* __restrict__ pts, StateArgs p,
                                                          const float4* __restrict__ PPs, float4* __restrict__ diag) {
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (size_t)p.n_scan * p.nz) return;
    const int ip = (int)(s / p.nz);
    const int iz = (int)(s - (size_t)ip * p.nz);
    const f3 tip = tip_anchor(pts, p, ip, iz);
    const float4 PP = PPs[s];
    int status = 0;
    StateEval e;
    eval_state(g, p, tip, PP.x, PP.y, PP.z, e);
    const double res = norm3(e.G);
    if (!(PP.w > 0.f) || !(res < kResidualGate)) status |= 1;
    double cof[3][3], det;
    double cond = INFINITY;
    if (cofactors(e.J, cof, det)) cond = cond_inf(e.J, cof, det);
    if (!(cond < kCondGate)) status |= 4;
    double S[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) S[i][j] = -0.5 * (e.J[i][j] + e.J[j][i]);
    const double lmin = eigmin_sym3(S);
    if (!(lmin > kStableGate)) status |= 8;
    if (!(isfinite(PP.x) && isfinite(PP.y) && isfinite(PP.z) && isfinite(res))) status |= 2;
    diag[s] = make_float4((float)res, (float)lmin, (float)cond, (float)status);
}

GridDev make_grid(const float* grid, const PpmGridSpec& s) {
    GridDev g;
    g.data = reinterpret_cast<const float4*>(grid);
    g.nx = s.nx; g.ny = s.ny; g.nz = s.nz;
    for (int q = 0; q < 4; ++q) { g.da[q] = s.dinvA[q]; g.db[q] = s.dinvB[q]; g.dc[q] = s.dinvC[q]; }
    g.klo = 0;
    g.nzw = s.nz;
    g.flag = nullptr;
    return g;
}

int n_waves(int n_scan, int scan_ny) {
    if (scan_ny > 0) {
        const int nxs = n_scan / scan_ny;
        return ((nxs + 7) / 8) * ((scan_ny + 7) / 8);
    }
    return (n_scan + 63) / 64;
}

}  // namespace

int launch_relax_tilted(const float* grid, const PpmGridSpec& s, const float* pts, int n_scan, int nz,
                        const PpmTip& tip, const PpmRelaxCtl& ctl, float* FEs, float* PPs, hipStream_t stream,
                        int klo, int nzw, int* flag, const float* dfW, const int* dfStart, int taps, int nout,
                        float* df) {
    TiltArgs p{};
    p.n_scan = n_scan; p.nz = nz; p.max_iter = ctl.max_iter; p.check_before = ctl.check_before;
    p.scan_ny = ctl.scan_ny; p.f2conv = ctl.f2conv; p.damped = ctl.damped != 0;
    for (int q = 0; q < 3; ++q) { p.A[q] = tip.tipA[q]; p.B[q] = tip.tipB[q]; p.C[q] = tip.tipC[q]; }
    p.dtip = tip.tipC[3];
    for (int q = 0; q < 4; ++q) { p.k[q] = tip.stiffness[q]; p.d0[q] = tip.dpos0[q]; }
    p.dt0 = tip.relax[0]; p.damp0 = tip.relax[1]; p.surf = tip.surfFF[0];
    const int G = ctl.lanes > 0 ? ctl.lanes : relax_group(n_scan);
    const dim3 blocks((unsigned)group_waves(n_scan, ctl.scan_ny, G));
    GridDev gd = make_grid(grid, s);
    if (nzw > 0) { gd.klo = klo; gd.nzw = nzw; gd.flag = flag; }
    const float4* P = reinterpret_cast<const float4*>(pts);
    float4* F = reinterpret_cast<float4*>(FEs);
    float4* Q = reinterpret_cast<float4*>(PPs);
    // LDS per workgroup: the Fz columns of the wave (nz x 64 floats) up to 64 KiB, else the df reads
    // the stored columns back; for G = 1 also the staged output runs, when the launch's waves stay
    // resident with them (at most as many per compute unit as the LDS then holds: the whole
    // 65,536-point flake scan is 4 per unit), since a latency-bound relaxation needs its waves more
    // than the instructions it saves
    const size_t fz = (size_t)nz * 64 * sizeof(float);
    const bool fz_lds = df && fz <= 65536;
    const size_t lds = fz_lds ? fz : 0;
    const long long cus = device_cus();
    const bool stage = G == 1 && lds + kStageBytes <= 65536 &&
                       ((long long)blocks.x + cus - 1) / cus <= device_lds_per_cu() / (long long)(lds + kStageBytes);
    DfOut dfo{dfW, dfStart, taps, nout, fz_lds};
    const bool fast = fast_frame(p, gd);
#define PPM_TILT(GG, FF, SS, L) \
    hipLaunchKernelGGL((k_relax_tilted<GG, FF, SS>), blocks, dim3(64), L, stream, gd, P, p, F, Q, dfo, df)
    if (G == 4) { if (fast) PPM_TILT(4, true, false, lds); else PPM_TILT(4, false, false, lds); }
    else if (G == 2) { if (fast) PPM_TILT(2, true, false, lds); else PPM_TILT(2, false, false, lds); }
    else if (stage) {
        if (fast) PPM_TILT(1, true, true, lds + kStageBytes);
        else PPM_TILT(1, false, true, lds + kStageBytes);
    }
    else { if (fast) PPM_TILT(1, true, false, lds); else PPM_TILT(1, false, false, lds); }
#undef PPM_TILT
    return check_launch("relax_strokes_tilted");
}

int launch_relax_strokes(const float* grid, const PpmGridSpec& s, const float* pts, int n_scan, int nz,
                         const float dTip[4], const float stiffness[4], const float dpos0[4], const float relax[4],
                         const PpmRelaxCtl& ctl, float* FEs, float* disps, float* PPs, hipStream_t stream) {
    StrokeArgs p{};
    p.n_scan = n_scan; p.nz = nz; p.max_iter = ctl.max_iter; p.scan_ny = ctl.scan_ny; p.f2conv = ctl.f2conv;
    p.damped = ctl.damped != 0;
    for (int q = 0; q < 3; ++q) p.dTip[q] = dTip[q];
    for (int q = 0; q < 4; ++q) { p.k[q] = stiffness[q]; p.d0[q] = dpos0[q]; }
    p.dt0 = relax[0]; p.damp0 = relax[1];
    const int G = ctl.lanes > 0 ? ctl.lanes : relax_group(n_scan);
    const dim3 blocks((unsigned)group_waves(n_scan, ctl.scan_ny, G));
    const GridDev gd = make_grid(grid, s);
    const float4* P = reinterpret_cast<const float4*>(pts);
    float4 *F = reinterpret_cast<float4*>(FEs), *Dp = reinterpret_cast<float4*>(disps),
           *Q = reinterpret_cast<float4*>(PPs);
    if (G == 4) hipLaunchKernelGGL(k_relax_strokes<4>, blocks, dim3(64), 0, stream, gd, P, p, F, Dp, Q);
    else if (G == 2) hipLaunchKernelGGL(k_relax_strokes<2>, blocks, dim3(64), 0, stream, gd, P, p, F, Dp, Q);
    else hipLaunchKernelGGL(k_relax_strokes<1>, blocks, dim3(64), 0, stream, gd, P, p, F, Dp, Q);
    return check_launch("relax_strokes");
}

int launch_relax_strokes_2d(const float* grid, const PpmGridSpec& s, const float* pts, int n_scan, int nz,
                            float k_lat, float dh, float dt, float damp, const PpmRelaxCtl& ctl, float* FEs,
                            float* disps, float* PPs, hipStream_t stream) {
    Strokes2DArgs p{};
    p.n_scan = n_scan; p.nz = nz; p.max_iter = ctl.max_iter; p.scan_ny = ctl.scan_ny; p.f2conv = ctl.f2conv;
    p.k_lat = k_lat; p.dh = dh; p.dt = dt; p.damp = damp;
    const int G = ctl.lanes > 0 ? ctl.lanes : relax_group(n_scan);
    const dim3 blocks((unsigned)group_waves(n_scan, ctl.scan_ny, G));
    const GridDev gd = make_grid(grid, s);
    const float4* P = reinterpret_cast<const float4*>(pts);
    float4 *F = reinterpret_cast<float4*>(FEs), *Dp = reinterpret_cast<float4*>(disps),
           *Q = reinterpret_cast<float4*>(PPs);
    if (G == 4) hipLaunchKernelGGL(k_relax_strokes_2d<4>, blocks, dim3(64), 0, stream, gd, P, p, F, Dp, Q);
    else if (G == 2) hipLaunchKernelGGL(k_relax_strokes_2d<2>, blocks, dim3(64), 0, stream, gd, P, p, F, Dp, Q);
    else hipLaunchKernelGGL(k_relax_strokes_2d<1>, blocks, dim3(64), 0, stream, gd, P, p, F, Dp, Q);
    return check_launch("relax_strokes_2d");
}

static StateArgs state_args(int n_scan, int nz, const PpmTip& tip, int steps) {
    StateArgs p{};
    p.n_scan = n_scan; p.nz = nz; p.steps = steps;
    for (int q = 0; q < 3; ++q) { p.A[q] = tip.tipA[q]; p.B[q] = tip.tipB[q]; p.C[q] = tip.tipC[q]; }
    p.dtip = tip.tipC[3];
    for (int q = 0; q < 4; ++q) { p.k[q] = tip.stiffness[q]; p.d0[q] = tip.dpos0[q]; }
    p.surf = tip.surfFF[0];
    return p;
}

int launch_relax_adjoint(const float* grid, const PpmGridSpec& s, const float* pts, int n_scan, int nz,
                         const PpmTip& tip, const float* PPs, const float* dL_dFEs, float* grid_cot, float* lambdas,
                         hipStream_t stream) {
    const size_t n = (size_t)n_scan * nz;
    hipLaunchKernelGGL(k_relax_adjoint, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, make_grid(grid, s),
                       reinterpret_cast<const float4*>(pts), state_args(n_scan, nz, tip, 0),
                       reinterpret_cast<const float4*>(PPs), reinterpret_cast<const float4*>(dL_dFEs), grid_cot,
                       reinterpret_cast<float4*>(lambdas));
    return check_launch("relax_adjoint");
}

int launch_relax_state_diag(const float* grid, const PpmGridSpec& s, const float* pts, int n_scan, int nz,
                            const PpmTip& tip, const float* PPs, float* diag, hipStream_t stream) {
    const size_t n = (size_t)n_scan * nz;
    hipLaunchKernelGGL(k_relax_state_diag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       make_grid(grid, s), reinterpret_cast<const float4*>(pts), state_args(n_scan, nz, tip, 0),
                       reinterpret_cast<const float4*>(PPs), reinterpret_cast<float4*>(diag));
    return check_launch("relax_state_diag");
}

int launch_relax_polish(const float* grid, const PpmGridSpec& s, const float* pts, int n_scan, int nz,
                        const PpmTip& tip, int steps, float* PPs, float* FEs, hipStream_t stream) {
    const size_t n = (size_t)n_scan * nz;
    hipLaunchKernelGGL(k_relax_polish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, make_grid(grid, s),
                       reinterpret_cast<const float4*>(pts), state_args(n_scan, nz, tip, steps),
                       reinterpret_cast<float4*>(PPs), reinterpret_cast<float4*>(FEs));
    return check_launch("relax_polish");
}

int launch_fe_in_strokes(const float* grid, const PpmGridSpec& s, const float* pts, int n_scan, int nz,
                         const PpmTip& tip, const float dTip[4], const float dpos0[4], float* FEs,
                         hipStream_t stream) {
    RawArgs p{};
    p.n_scan = n_scan; p.nz = nz; p.scan_ny = 0;
    for (int q = 0; q < 3; ++q) {
        p.A[q] = tip.tipA[q]; p.B[q] = tip.tipB[q]; p.C[q] = tip.tipC[q];
        p.dTip[q] = dTip[q]; p.d0[q] = dpos0[q];
    }
    hipLaunchKernelGGL(k_fe_in_strokes, dim3((n_scan + 63) / 64), dim3(64), 0, stream, make_grid(grid, s),
                       reinterpret_cast<const float4*>(pts), p, reinterpret_cast<float4*>(FEs));
    return check_launch("fe_in_strokes_tilted");
}

}  // namespace ppm
