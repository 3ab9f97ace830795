// ppm_oracle.cpp — CPU oracle for the probe-particle AFM path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (spammm_amd/) links, loads
// or calls this file.  It is used by tests/ (as the checker), by
// __graft_entry__.smoke() (as the checker) and by bench.py's cpu_baseline leg
// (as the timed CPU restatement of the reference's algorithm).
//
// It restates, from the reference's source text, the semantics of:
//   * the pair law + point-charge Coulomb of evalLJC_QZs_toImg /
//     evalMorseC_QZs_toImg  (kernels/AFM.cl:730-757, :916-991, :996-1056),
//     including the |E|>100 rescale (kernels/AFM.cl:984-985, :1053-1054),
//     with the periodic-image convention of make_MorseFF
//     (kernels/gridFF.cl:938-996: loop over ix,iy in [-nPBC,nPBC]);
//   * trilinear sampling with normalized coordinates, REPEAT addressing and
//     linear filtering (sampler_1, kernels/AFM.cl:49, :100-105), written as the
//     OpenCL 1.2 specification's linear-filter REPEAT formula, combined in the
//     nested order of read_imagef_trilin_ (kernels/AFM.cl:81-97);
//   * tipForce (kernels/AFM.cl:53-58), update_FIRE (kernels/AFM.cl:121-157),
//     rotMat/rotMatT (kernels/common.cl:44-45);
//   * relaxStrokesTilted (kernels/AFM.cl:478-562), relaxStrokes
//     (kernels/AFM.cl:345-394) and getFEinStrokesTilted (kernels/AFM.cl:210-233);
//   * the gridless direct Morse path: relaxStrokesTiltedMorseDirect
//     (kernels/AFM.cl:1419-1511, pair sum evalMorseCDirect_local :1391-1407) and its
//     implicit-equilibrium adjoint morseDirectStateAdjoint / morseDirectParamPartials /
//     reduceMorseDirectParamPartials (kernels/AFM.cl:1546-2029), restated in double.
//
// Each function exists in two precisions:
//   *_f64 — every operation in double: the error oracle;
//   *_f32 — every operation in float, in a fixed operation order, built with
//           -ffp-contract=off: the FP32 op-order-faithful restatement.
//
// Parity status of each function is recorded in the header comment of the
// Python wrapper (oracle/oracle.py).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>

namespace {

constexpr double kR2SAFE = 1e-4;                 // kernels/common.cl:28
constexpr double kCOULOMB = 14.3996448915;       // kernels/common.cl:29

template <class R> struct V3 {
    R x, y, z;
};
template <class R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> add(V3<R> a, V3<R> b) { return v3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> inline V3<R> sub(V3<R> a, V3<R> b) { return v3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> inline V3<R> mul(V3<R> a, R s) { return v3<R>(a.x * s, a.y * s, a.z * s); }
template <class R> inline V3<R> mulv(V3<R> a, V3<R> b) { return v3<R>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// rotMat(v,a,b,c) = (dot(v,a),dot(v,b),dot(v,c))    kernels/common.cl:44
template <class R> inline V3<R> rotMat(V3<R> v, V3<R> a, V3<R> b, V3<R> c) {
    return v3<R>(dot(v, a), dot(v, b), dot(v, c));
}
// rotMatT(v,a,b,c) = a*v.x + b*v.y + c*v.z          kernels/common.cl:45
template <class R> inline V3<R> rotMatT(V3<R> v, V3<R> a, V3<R> b, V3<R> c) {
    return add(add(mul(a, v.x), mul(b, v.y)), mul(c, v.z));
}
template <class R, class P> inline V3<R> ld3(const P* p) { return v3<R>((R)p[0], (R)p[1], (R)p[2]); }

// tipForce   kernels/AFM.cl:53-58
template <class R, class P> inline V3<R> tipForce(V3<R> dpos, const P* k, const P* d0) {
    R r = std::sqrt(dot(dpos, dpos));
    r = std::max(r, (R)1e-10f);
    V3<R> a = mulv(sub(dpos, ld3<R>(d0)), ld3<R>(k));
    R s = ((R)k[3] * (r - (R)d0[3])) / r;
    return add(a, mul(dpos, s));
}

// Grid + sampling description shared by the relax functions.
template <class GT> struct GridT {
    const GT* data;      // (nz,ny,nx,4), x fastest
    int n[3];
    double dinv[3][4];   // world -> normalized coordinate rows (.w carries -origin/L + 0.5/n)
};
using Grid = GridT<float>;

// One axis of the OpenCL linear-filter REPEAT addressing (OpenCL 1.2 spec, 8.2):
//   u = (s - floor(s))*n;  i0 = floor(u-0.5);  a = frac(u-0.5);  i1 = i0+1;
//   if(i0<0) i0+=n;  if(i1>n-1) i1-=n.
template <class R> inline void axis_repeat(R s, int n, int& i0, int& i1, R& a) {
    R u = (s - std::floor(s)) * (R)n;
    R t = u - (R)0.5f;
    R fl = std::floor(t);
    a = t - fl;
    fl = std::fmin(std::fmax(fl, (R)-1), (R)n);  // keeps non-finite coordinates in range
    int i = (int)fl;
    int j = i + 1;
    if (i < 0) i += n;
    if (j > n - 1) j -= n;
    // non-finite coordinates: keep the index in range (value is garbage anyway)
    i = std::min(std::max(i, 0), n - 1);
    j = std::min(std::max(j, 0), n - 1);
    i0 = i; i1 = j;
}

template <class R> struct FE { R x, y, z, w; };

// interpFE  kernels/AFM.cl:100-105 (coordinate = dot((pos,1), dinv)); trilinear in the
// nested order of read_imagef_trilin_  kernels/AFM.cl:81-97.
template <class R, class GT> inline FE<R> interp(const GridT<GT>& g, V3<R> p) {
    R s[3];
    for (int k = 0; k < 3; ++k)
        s[k] = p.x * (R)g.dinv[k][0] + p.y * (R)g.dinv[k][1] + p.z * (R)g.dinv[k][2] + (R)g.dinv[k][3];
    int i0, i1, j0, j1, k0, k1;
    R a, b, c;
    axis_repeat<R>(s[0], g.n[0], i0, i1, a);
    axis_repeat<R>(s[1], g.n[1], j0, j1, b);
    axis_repeat<R>(s[2], g.n[2], k0, k1, c);
    const R ma = (R)1.0f - a, mb = (R)1.0f - b, mc = (R)1.0f - c;
    const int nx = g.n[0], ny = g.n[1];
    auto T = [&](int i, int j, int k) -> const GT* { return g.data + 4 * (((size_t)k * ny + j) * nx + i); };
    const GT *c000 = T(i0, j0, k0), *c100 = T(i1, j0, k0), *c010 = T(i0, j1, k0), *c110 = T(i1, j1, k0);
    const GT *c001 = T(i0, j0, k1), *c101 = T(i1, j0, k1), *c011 = T(i0, j1, k1), *c111 = T(i1, j1, k1);
    R o[4];
    for (int q = 0; q < 4; ++q) {
        R lo = ((R)c000[q] * ma + (R)c100[q] * a) * mb + ((R)c010[q] * ma + (R)c110[q] * a) * b;
        R hi = ((R)c001[q] * ma + (R)c101[q] * a) * mb + ((R)c011[q] * ma + (R)c111[q] * a) * b;
        o[q] = lo * mc + hi * c;
    }
    return FE<R>{o[0], o[1], o[2], o[3]};
}

// update_FIRE   kernels/AFM.cl:121-157 (OPT_FIRE=1)
template <class R> inline V3<R> update_FIRE(V3<R> f, V3<R> v, R& dt, R& damp, R dtmin, R dtmax, R damp0) {
    R ff = dot(f, f);
    R vv = dot(v, v);
    R vf = dot(v, f);
    if (vf < (R)0) {
        v = v3<R>(0, 0, 0);
        dt = std::max(dtmin, dt * (R)0.5f);
        damp = damp0;
    } else {
        v = mul(v, (R)1.0f - damp);
        v = add(v, mul(f, damp * std::sqrt(vv / (ff + (R)1e-8f))));
        dt = std::min(dtmax, dt * (R)1.1f);
        damp = damp * (R)0.99f;
    }
    return v;
}

// The integrator step before `v += f dt; pos += v dt`: update_FIRE (OPT_FIRE=1), or the damped
// velocity `v *= (1 - damp)` of the OPT_FIRE=0 build (kernels/AFM.cl:378-382, :536-540), in which
// dt and damp keep their initial values.
template <class R> inline V3<R> integrate(int damped, V3<R> f, V3<R> v, R& dt, R& damp, R dtmin, R dtmax, R damp0) {
    if (damped) return mul(v, (R)1.0f - damp);
    return update_FIRE<R>(f, v, dt, damp, dtmin, dtmax, damp0);
}

// relaxStrokesTilted   kernels/AFM.cl:478-562, plus PP telemetry (pos, +iters on
// convergence / -iters otherwise) in the convention of relaxStrokesTiltedMorseDirect
// (kernels/AFM.cl:1494-1505).  check_before=1 moves the |f|^2 test before the update
// as relaxStrokesTiltedMorseDirect does (kernels/AFM.cl:1478).
// `sample(pos)` returns the sample FE at a world point: the trilinear grid sample for
// relaxStrokesTilted, the direct pair sum for relaxStrokesTiltedMorseDirect.
template <class R, class SRC, class P>
void relax_tilted_src(const SRC& sample, const P* pts, int n_scan, int nz, const P* tipA, const P* tipB,
                      const P* tipC, const P* stiffness, const P* dpos0, const P* relax, const P* surfFF, int max_iter,
                      double f2conv_f, int check_before, P* FEs, P* PPs, int damped = 0, int f32_anchors = 0) {
    const V3<R> A = ld3<R>(tipA), B = ld3<R>(tipB), C = ld3<R>(tipC);
    const R f2conv = (R)f2conv_f;
#pragma omp parallel for schedule(dynamic, 8)
    for (int ip = 0; ip < n_scan; ++ip) {
        const V3<R> dTip = mul(C, (R)tipC[3]);
        const V3<R> d0r = rotMatT(ld3<R>(dpos0), A, B, C);
        V3<R> tipPos = ld3<R>(pts + 4 * (size_t)ip);
        // f32_anchors: the anchor advanced as the FP32 kernel advances it (float32 position plus the
        // float32 step tipC.xyz * tipC.w, kernels/AFM.cl:556), so an FP64 relaxation runs at the FP32
        // scan's exact heights
        V3<float> tipf = v3<float>((float)pts[4 * (size_t)ip], (float)pts[4 * (size_t)ip + 1],
                                   (float)pts[4 * (size_t)ip + 2]);
        const V3<float> dTipf = mul(v3<float>((float)tipC[0], (float)tipC[1], (float)tipC[2]), (float)tipC[3]);
        if (f32_anchors) tipPos = v3<R>((R)tipf.x, (R)tipf.y, (R)tipf.z);
        V3<R> pos = add(tipPos, d0r);
        R dt = (R)relax[0], damp = (R)relax[1];
        const R dtmax = dt, dtmin = dtmax * (R)0.1f, damp0 = damp;
        for (int iz = 0; iz < nz; ++iz) {
            V3<R> v = v3<R>(0, 0, 0);
            int it = 0;
            bool conv = false;
            for (it = 0; it < max_iter; ++it) {
                FE<R> fe = sample(pos);
                V3<R> f = v3<R>(fe.x, fe.y, fe.z);
                V3<R> dpos = sub(pos, tipPos);
                V3<R> dpos_ = rotMat(dpos, A, B, C);
                V3<R> ftip = tipForce<R>(dpos_, stiffness, dpos0);
                f = add(f, rotMatT(ftip, A, B, C));
                f = add(f, mul(C, (R)surfFF[0]));
                R ff = dot(f, f);
                if (check_before && ff < f2conv) { conv = true; break; }
                v = integrate<R>(damped, f, v, dt, damp, dtmin, dtmax, damp0);
                v = add(v, mul(f, dt));
                pos = add(pos, mul(v, dt));
                if (!check_before && ff < f2conv) { conv = true; break; }
            }
            int iters = conv ? it + 1 : max_iter;
            FE<R> fe = sample(pos);
            V3<R> fr = rotMat(v3<R>(fe.x, fe.y, fe.z), A, B, C);
            const size_t o = 4 * ((size_t)ip * nz + iz);
            FEs[o + 0] = (P)fr.x; FEs[o + 1] = (P)fr.y; FEs[o + 2] = (P)fr.z; FEs[o + 3] = (P)fe.w;
            bool finite = std::isfinite(fe.x) && std::isfinite(fe.y) && std::isfinite(fe.z) && std::isfinite(fe.w) &&
                          std::isfinite(pos.x) && std::isfinite(pos.y) && std::isfinite(pos.z);
            if (PPs) {
                PPs[o + 0] = (P)pos.x; PPs[o + 1] = (P)pos.y; PPs[o + 2] = (P)pos.z;
                PPs[o + 3] = (P)((conv && finite) ? iters : -iters);
            }
            tipPos = add(tipPos, dTip);
            if (f32_anchors) {
                tipf = add(tipf, dTipf);
                tipPos = v3<R>((R)tipf.x, (R)tipf.y, (R)tipf.z);
            }
            pos = add(pos, dTip);
        }
    }
}

template <class R, class GT, class P>
void relax_tilted(const GridT<GT>& g, const P* pts, int n_scan, int nz, const P* tipA, const P* tipB, const P* tipC,
                  const P* stiffness, const P* dpos0, const P* relax, const P* surfFF, int max_iter, double f2conv_f,
                  int check_before, P* FEs, P* PPs, int damped = 0, int f32_anchors = 0) {
    relax_tilted_src<R>([&g](V3<R> p) { return interp<R>(g, p); }, pts, n_scan, nz, tipA, tipB, tipC, stiffness,
                        dpos0, relax, surfFF, max_iter, f2conv_f, check_before, FEs, PPs, damped, f32_anchors);
}

// relaxStrokes   kernels/AFM.cl:345-394.  FE output is the sample taken at the start
// of the last iteration (before the final update); disps = pos - (tipPos + dpos0).
template <class R, class GT, class P>
void relax_strokes(const GridT<GT>& g, const P* pts, int n_scan, int nz, const P* dTip_, const P* stiffness,
                   const P* dpos0, const P* relax, int max_iter, double f2conv_f, P* FEs, P* disps, P* PPs,
                   int damped = 0) {
    const R f2conv = (R)f2conv_f;
#pragma omp parallel for schedule(dynamic, 8)
    for (int ip = 0; ip < n_scan; ++ip) {
        const V3<R> dTip = ld3<R>(dTip_);
        const V3<R> d0 = ld3<R>(dpos0);
        V3<R> tipPos = ld3<R>(pts + 4 * (size_t)ip);
        V3<R> pos = add(tipPos, d0);
        R dt = (R)relax[0], damp = (R)relax[1];
        const R dtmax = dt, dtmin = dtmax * (R)0.1f, damp0 = damp;
        for (int iz = 0; iz < nz; ++iz) {
            FE<R> fe{0, 0, 0, 0};
            V3<R> v = v3<R>(0, 0, 0);
            int it = 0;
            bool conv = false;
            for (it = 0; it < max_iter; ++it) {
                fe = interp<R>(g, pos);
                V3<R> f = v3<R>(fe.x, fe.y, fe.z);
                f = add(f, tipForce<R>(sub(pos, tipPos), stiffness, dpos0));
                R ff = dot(f, f);
                v = integrate<R>(damped, f, v, dt, damp, dtmin, dtmax, damp0);
                v = add(v, mul(f, dt));
                pos = add(pos, mul(v, dt));
                if (ff < f2conv) { conv = true; break; }
            }
            int iters = conv ? it + 1 : max_iter;
            const size_t o = 4 * ((size_t)ip * nz + iz);
            FEs[o + 0] = (P)fe.x; FEs[o + 1] = (P)fe.y; FEs[o + 2] = (P)fe.z; FEs[o + 3] = (P)fe.w;
            V3<R> dd = sub(pos, add(tipPos, d0));
            if (disps) {
                disps[o + 0] = (P)dd.x; disps[o + 1] = (P)dd.y; disps[o + 2] = (P)dd.z; disps[o + 3] = 0.f;
            }
            bool finite = std::isfinite(fe.x) && std::isfinite(fe.y) && std::isfinite(fe.z) && std::isfinite(fe.w) &&
                          std::isfinite(pos.x) && std::isfinite(pos.y) && std::isfinite(pos.z);
            if (PPs) {
                PPs[o + 0] = (P)pos.x; PPs[o + 1] = (P)pos.y; PPs[o + 2] = (P)pos.z;
                PPs[o + 3] = (P)((conv && finite) ? iters : -iters);
            }
            tipPos = add(tipPos, dTip);
            pos = add(pos, dTip);
        }
    }
}

// relaxStrokes2D   kernels/AFM.cl:396-450: z fixed per slice at az - iz*dh, x,y relaxed from the
// anchor by damped velocity steps in the grid force plus -k_lat (p - anchor); the |F_xy|^2 test
// follows the update.  FEs = sample at the final position, disps = (x - ax, y - ay, 0, 0), PPs =
// (x, y, pz, +/-iterations).
template <class R, class GT, class P>
void relax_2d(const GridT<GT>& g, const P* pts, int n_scan, int nz, double k_lat_f, double dh_f, double dt_f,
              double damp_f, int max_iter, double f2conv_f, P* FEs, P* disps, P* PPs) {
    const R k_lat = (R)k_lat_f, dh = (R)dh_f, dt = (R)dt_f, damp = (R)damp_f, f2conv = (R)f2conv_f;
#pragma omp parallel for schedule(dynamic, 8)
    for (int ip = 0; ip < n_scan; ++ip) {
        const R ax = (R)pts[4 * (size_t)ip], ay = (R)pts[4 * (size_t)ip + 1], az = (R)pts[4 * (size_t)ip + 2];
        for (int iz = 0; iz < nz; ++iz) {
            const R pz = az - (R)iz * dh;
            R px = ax, py = ay, vx = 0, vy = 0;
            int it = 0;
            bool conv = false;
            for (it = 0; it < max_iter; ++it) {
                FE<R> fe = interp<R>(g, v3<R>(px, py, pz));
                const R fx = fe.x - k_lat * (px - ax);
                const R fy = fe.y - k_lat * (py - ay);
                vx = vx * ((R)1.0f - damp) + fx * dt;
                vy = vy * ((R)1.0f - damp) + fy * dt;
                px += vx * dt;
                py += vy * dt;
                if ((fx * fx + fy * fy) < f2conv) { conv = true; break; }
            }
            const int iters = conv ? it + 1 : max_iter;
            FE<R> fe = interp<R>(g, v3<R>(px, py, pz));
            const size_t o = 4 * ((size_t)ip * nz + iz);
            FEs[o + 0] = (P)fe.x; FEs[o + 1] = (P)fe.y; FEs[o + 2] = (P)fe.z; FEs[o + 3] = (P)fe.w;
            if (disps) { disps[o + 0] = (P)(px - ax); disps[o + 1] = (P)(py - ay); disps[o + 2] = 0; disps[o + 3] = 0; }
            const bool finite = std::isfinite(fe.x) && std::isfinite(fe.y) && std::isfinite(fe.z) &&
                                std::isfinite(fe.w) && std::isfinite(px) && std::isfinite(py);
            if (PPs) {
                PPs[o + 0] = (P)px; PPs[o + 1] = (P)py; PPs[o + 2] = (P)pz;
                PPs[o + 3] = (P)((conv && finite) ? iters : -iters);
            }
        }
    }
}

// getFEinStrokesTilted   kernels/AFM.cl:210-233 (dpos0 is NOT rotated here).
template <class R, class GT, class P>
void fe_in_strokes_tilted(const GridT<GT>& g, const P* pts, int n_scan, int nz, const P* tipA, const P* tipB,
                          const P* tipC, const P* dTip_, const P* dpos0, P* FEs) {
    const V3<R> A = ld3<R>(tipA), B = ld3<R>(tipB), C = ld3<R>(tipC);
#pragma omp parallel for schedule(static)
    for (int ip = 0; ip < n_scan; ++ip) {
        V3<R> pos = add(ld3<R>(pts + 4 * (size_t)ip), ld3<R>(dpos0));
        const V3<R> dTip = ld3<R>(dTip_);
        for (int iz = 0; iz < nz; ++iz) {
            FE<R> fe = interp<R>(g, pos);
            V3<R> fr = rotMat(v3<R>(fe.x, fe.y, fe.z), A, B, C);
            const size_t o = 4 * ((size_t)ip * nz + iz);
            FEs[o + 0] = (P)fr.x; FEs[o + 1] = (P)fr.y; FEs[o + 2] = (P)fr.z; FEs[o + 3] = (P)fe.w;
            pos = add(pos, dTip);
        }
    }
}

// Pair law of one atom (plus its Coulomb sites) at world point p.
//   getLJ          kernels/AFM.cl:738-745 : ir2=1/(r2+R2SAFE); ir6=ir2^3; E=(c12*ir6-c6)*ir6;
//                                            F=((12*c12*ir6-6*c6)*ir6*ir2)*dp
//   getMorse       kernels/AFM.cl:747-757 : r=sqrt(r2+R2SAFE); s=exp(K*(r-R0)); E=E0*s*(s-2);
//                                            F=dp*(-E0*s*(s-1)*2*K/r)
//   getCoulombAFM  kernels/AFM.cl:730-736 : ir2=1/(r2+R2SAFE); E=q*sqrt(ir2); F=dp*(E*ir2)
// model 0 = LJ (cpar = c6,c12), model 1 = Morse (cpar = R0,E0,K).
template <class R>
inline void pair_atom(int model, V3<R> p, V3<R> a, R q, const R* cp, const R* Qs, const R* QZs, R acc[4]) {
    V3<R> dp = sub(p, a);
    if (model == 0) {
        R ir2 = (R)1 / (dot(dp, dp) + (R)kR2SAFE);
        R ir6 = ir2 * ir2 * ir2;
        R E = (cp[1] * ir6 - cp[0]) * ir6;
        R fr = ((R)12 * cp[1] * ir6 - (R)6 * cp[0]) * ir6 * ir2;
        acc[0] += dp.x * fr; acc[1] += dp.y * fr; acc[2] += dp.z * fr; acc[3] += E;
    } else {
        R r = std::sqrt(dot(dp, dp) + (R)kR2SAFE);
        R ex = std::exp(cp[2] * (r - cp[0]));
        R E = cp[1] * ex * (ex - (R)2);
        R fr = -cp[1] * ex * (ex - (R)1) * (R)2 * cp[2];
        acc[0] += dp.x * (fr / r); acc[1] += dp.y * (fr / r); acc[2] += dp.z * (fr / r); acc[3] += E;
    }
    for (int k = 0; k < 4; ++k) {
        if (Qs[k] == (R)0) continue;   // a zero-weight site adds exactly zero
        V3<R> dk = sub(add(p, v3<R>(0, 0, QZs[k])), a);
        R ir2 = (R)1 / (dot(dk, dk) + (R)kR2SAFE);
        R E = q * std::sqrt(ir2);
        R e2 = E * ir2;
        acc[0] += (dk.x * e2) * Qs[k]; acc[1] += (dk.y * e2) * Qs[k]; acc[2] += (dk.z * e2) * Qs[k];
        acc[3] += E * Qs[k];
    }
}

template <class P> struct AtomsT {
    int model;
    int na;
    const P* atoms;  // na*4 (x,y,z,q)
    const P* cpar;   // na*cstride
    int cstride;
    double Qs[4], QZs[4];
    int npbc[2];
    double lvA[3], lvB[3];
};
using Atoms = AtomsT<float>;

// Total sample FE at world point p: sum over atoms, periodic images
// (ix in [-npbc0,npbc0], iy in [-npbc1,npbc1], atom shifted by ix*lvA + iy*lvB).
template <class R, class P> inline void sample_fe(const AtomsT<P>& S, V3<R> p, R acc[4]) {
    R Qs[4], QZs[4];
    for (int k = 0; k < 4; ++k) { Qs[k] = (R)S.Qs[k] * (R)kCOULOMB; QZs[k] = (R)S.QZs[k]; }
    acc[0] = acc[1] = acc[2] = acc[3] = 0;
    for (int j = 0; j < S.na; ++j) {
        const P* at = S.atoms + 4 * (size_t)j;
        R cp[4] = {0, 0, 0, 0};
        for (int c = 0; c < S.cstride && c < 4; ++c) cp[c] = (R)S.cpar[(size_t)j * S.cstride + c];
        for (int iy = -S.npbc[1]; iy <= S.npbc[1]; ++iy)
            for (int ix = -S.npbc[0]; ix <= S.npbc[0]; ++ix) {
                V3<R> a = v3<R>((R)at[0] + (R)ix * (R)S.lvA[0] + (R)iy * (R)S.lvB[0],
                                (R)at[1] + (R)ix * (R)S.lvA[1] + (R)iy * (R)S.lvB[1],
                                (R)at[2] + (R)ix * (R)S.lvA[2] + (R)iy * (R)S.lvB[2]);
                pair_atom<R>(S.model, p, a, (R)at[3], cp, Qs, QZs, acc);
            }
    }
}


// ---------------------------------------------------------------------------------------
// Gridless direct Morse path (aperiodic, uncapped): relaxStrokesTiltedMorseDirect
// kernels/AFM.cl:1419-1511 — relaxStrokesTilted with the direct pair sum as the sample and
// the |f|^2 test before the update (kernels/AFM.cl:1478).
template <class R, class P>
void relax_direct(const P* atoms, const P* cMs, int na, const P* Qs, const P* QZs, const P* pts, int n_scan, int nz,
                  const P* tipA, const P* tipB, const P* tipC, const P* stiffness, const P* dpos0, const P* relax,
                  const P* surfFF, int max_iter, double f2conv, P* FEs, P* PPs) {
    AtomsT<P> S{1, na, atoms, cMs, 4, {}, {}, {0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < 4; ++k) { S.Qs[k] = (double)Qs[k]; S.QZs[k] = (double)QZs[k]; }
    auto sample = [&S](V3<R> p) {
        R acc[4];
        sample_fe<R>(S, p, acc);
        return FE<R>{acc[0], acc[1], acc[2], acc[3]};
    };
    relax_tilted_src<R>(sample, pts, n_scan, nz, tipA, tipB, tipC, stiffness, dpos0, relax, surfFF, max_iter, f2conv,
                        1, FEs, PPs);
}

// Per-atom pair quantities of the direct Morse + Coulomb law at probe position q (double):
// force F_i on the probe, energy E_i, energy Hessian H_i = d2E_i/dq2 and the parameter
// derivatives used by the adjoint (kernels/AFM.cl:1546-1614 sample FE + Hessian,
// :1873-1940 partials).
struct PairD {
    double F[3], E, H[3][3];
    double dF_dR0[3], dF_dE0[3], dF_dQ[3];
    double dE_dR0, dE_dE0, dE_dQ;
};

inline void pair_direct(const double q[3], const double a[4], const double cm[4], const double Qk[4],
                        const double QZk[4], PairD& o) {
    const double d[3] = {q[0] - a[0], q[1] - a[1], q[2] - a[2]};
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + kR2SAFE;
    const double r = std::sqrt(r2);
    const double R0 = cm[0], E0 = cm[1], K = cm[2];
    const double s = std::exp(K * (r - R0));
    const double Ep = 2.0 * K * E0 * s * (s - 1.0);             // dE/dr
    const double Epp = 2.0 * K * K * E0 * s * (2.0 * s - 1.0);  // d2E/dr2
    const double cout = (Epp - Ep / r) / r2, cdia = Ep / r;
    o.E = E0 * s * (s - 2.0);
    for (int i = 0; i < 3; ++i) {
        o.F[i] = d[i] * (-Ep / r);
        o.dF_dR0[i] = d[i] * (2.0 * K * K * E0 * s * (2.0 * s - 1.0) / r);
        o.dF_dE0[i] = d[i] * (-2.0 * K * s * (s - 1.0) / r);
        o.dF_dQ[i] = 0.0;
        for (int j = 0; j < 3; ++j) o.H[i][j] = cout * d[i] * d[j] + (i == j ? cdia : 0.0);
    }
    o.dE_dR0 = -2.0 * K * E0 * s * (s - 1.0);
    o.dE_dE0 = s * (s - 2.0);
    o.dE_dQ = 0.0;
    for (int k = 0; k < 4; ++k) {
        if (Qk[k] == 0.0) continue;
        const double dk[3] = {q[0] - a[0], q[1] - a[1], q[2] + QZk[k] - a[2]};
        const double rk2 = dk[0] * dk[0] + dk[1] * dk[1] + dk[2] * dk[2] + kR2SAFE;
        const double rk = std::sqrt(rk2), rk3 = rk * rk2, rk5 = rk3 * rk2;
        const double C = a[3] * Qk[k];
        o.E += C / rk;
        o.dE_dQ += Qk[k] / rk;
        for (int i = 0; i < 3; ++i) {
            o.F[i] += dk[i] * (C / rk3);
            o.dF_dQ[i] += dk[i] * (Qk[k] / rk3);
            for (int j = 0; j < 3; ++j) o.H[i][j] += 3.0 * C * dk[i] * dk[j] / rk5 - (i == j ? C / rk3 : 0.0);
        }
    }
}

// Smallest eigenvalue of a symmetric 3x3 matrix (Smith, Commun. ACM 4(4):168, 1961).
inline double eigmin_sym3(const double S[3][3]) {
    const double p1 = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
    if (p1 == 0.0) return std::min(S[0][0], std::min(S[1][1], S[2][2]));
    const double q = (S[0][0] + S[1][1] + S[2][2]) / 3.0;
    const double a = S[0][0] - q, b = S[1][1] - q, c = S[2][2] - q;
    const double p = std::sqrt((a * a + b * b + c * c + 2.0 * p1) / 6.0);
    const double B00 = a / p, B11 = b / p, B22 = c / p, B01 = S[0][1] / p, B02 = S[0][2] / p, B12 = S[1][2] / p;
    const double detB = B00 * (B11 * B22 - B12 * B12) - B01 * (B01 * B22 - B12 * B02) + B02 * (B01 * B12 - B11 * B02);
    const double r = std::min(std::max(0.5 * detB, -1.0), 1.0);
    const double phi = std::acos(r) / 3.0;
    return q + 2.0 * p * std::cos(phi + 2.0 * M_PI / 3.0);
}

}  // namespace

extern "C" {

int oracle_version(void) { return 1; }

// Direct (brute-force) sample FE at arbitrary points, uncapped, in double
// (cs_brute_afm_morse_c_points semantics, kernels/contact_surface.cl:1341-1389,
//  extended to the LJ pair law and periodic images).
void oracle_pair_fe_f64(int model, const float* atoms, const float* cpar, int cstride, int na, const float* Qs,
                        const float* QZs, int npbc_a, int npbc_b, const float* lvA, const float* lvB,
                        const double* queries, int nq, double* out) {
    Atoms S{model, na, atoms, cpar, cstride, {}, {}, {npbc_a, npbc_b}, {}, {}};
    for (int k = 0; k < 4; ++k) { S.Qs[k] = Qs[k]; S.QZs[k] = QZs[k]; }
    for (int k = 0; k < 3; ++k) { S.lvA[k] = lvA[k]; S.lvB[k] = lvB[k]; }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nq; ++i) {
        double acc[4];
        sample_fe<double>(S, v3<double>(queries[3 * i], queries[3 * i + 1], queries[3 * i + 2]), acc);
        for (int q = 0; q < 4; ++q) out[4 * (size_t)i + q] = acc[q];
    }
}

// Same as oracle_pair_fe_f64 but with double-precision atom / parameter inputs, for
// finite-difference tests that perturb parameters by amounts below float resolution.
void oracle_pair_fe_f64d(int model, const double* atoms, const double* cpar, int cstride, int na, const double* Qs,
                         const double* QZs, int npbc_a, int npbc_b, const double* lvA, const double* lvB,
                         const double* queries, int nq, double* out) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nq; ++i) {
        double Qk[4], QZk[4];
        for (int k = 0; k < 4; ++k) { Qk[k] = Qs[k] * kCOULOMB; QZk[k] = QZs[k]; }
        double acc[4] = {0, 0, 0, 0};
        V3<double> p = v3<double>(queries[3 * i], queries[3 * i + 1], queries[3 * i + 2]);
        for (int j = 0; j < na; ++j) {
            double cp[4] = {0, 0, 0, 0};
            for (int c = 0; c < cstride && c < 4; ++c) cp[c] = cpar[(size_t)j * cstride + c];
            for (int iy = -npbc_b; iy <= npbc_b; ++iy)
                for (int ix = -npbc_a; ix <= npbc_a; ++ix) {
                    V3<double> a = v3<double>(atoms[4 * j] + ix * lvA[0] + iy * lvB[0],
                                              atoms[4 * j + 1] + ix * lvA[1] + iy * lvB[1],
                                              atoms[4 * j + 2] + ix * lvA[2] + iy * lvB[2]);
                    pair_atom<double>(model, p, a, atoms[4 * j + 3], cp, Qk, QZk, acc);
                }
        }
        for (int q = 0; q < 4; ++q) out[4 * (size_t)i + q] = acc[q];
    }
}

// Grid projection: voxel (ia,ib,ic) at p0 + dA*ia + dB*ib + dC*ic  (kernels/AFM.cl:959),
// output layout (nz,ny,nx,4), x fastest (write_imagef texel order, kernels/AFM.cl:989).
// cap!=0 applies: renorm=100/|E|; if(renorm<1) fe*=renorm  (kernels/AFM.cl:984-985).
#define GRID_PROJECT(NAME, R)                                                                                     \
    void NAME(int model, const float* atoms, const float* cpar, int cstride, int na, const float* Qs,              \
              const float* QZs, int npbc_a, int npbc_b, const float* lvA, const float* lvB, const int* n,          \
              const float* p0, const float* dA, const float* dB, const float* dC, int cap, float* out) {           \
        Atoms S{model, na, atoms, cpar, cstride, {}, {}, {npbc_a, npbc_b}, {}, {}};                                \
        for (int k = 0; k < 4; ++k) { S.Qs[k] = Qs[k]; S.QZs[k] = QZs[k]; }                                          \
        for (int k = 0; k < 3; ++k) { S.lvA[k] = lvA[k]; S.lvB[k] = lvB[k]; }                                        \
        const long nx = n[0], ny = n[1], nzz = n[2];                                                               \
        _Pragma("omp parallel for collapse(2) schedule(dynamic, 4)")                                                 \
        for (long ic = 0; ic < nzz; ++ic)                                                                          \
            for (long ib = 0; ib < ny; ++ib)                                                                       \
                for (long ia = 0; ia < nx; ++ia) {                                                                 \
                    V3<R> p = v3<R>((R)p0[0] + (R)dA[0] * (R)ia + (R)dB[0] * (R)ib + (R)dC[0] * (R)ic,              \
                                    (R)p0[1] + (R)dA[1] * (R)ia + (R)dB[1] * (R)ib + (R)dC[1] * (R)ic,              \
                                    (R)p0[2] + (R)dA[2] * (R)ia + (R)dB[2] * (R)ib + (R)dC[2] * (R)ic);             \
                    R acc[4];                                                                                      \
                    sample_fe<R>(S, p, acc);                                                                       \
                    if (cap) {                                                                                     \
                        R renorm = (R)100 / std::fabs(acc[3]);                                                     \
                        if (renorm < (R)1) for (int q = 0; q < 4; ++q) acc[q] *= renorm;                           \
                    }                                                                                              \
                    float* o = out + 4 * ((ic * ny + ib) * nx + ia);                                               \
                    for (int q = 0; q < 4; ++q) o[q] = (float)acc[q];                                              \
                }                                                                                                  \
    }
GRID_PROJECT(oracle_grid_project_f64, double)
GRID_PROJECT(oracle_grid_project_f32, float)

// Grid projection restricted to an explicit list of voxel indices (flat (ic*ny+ib)*nx+ia),
// double precision; used to check large grids on a sample of voxels.
void oracle_grid_voxels_f64(int model, const float* atoms, const float* cpar, int cstride, int na, const float* Qs,
                            const float* QZs, int npbc_a, int npbc_b, const float* lvA, const float* lvB, const int* n,
                            const float* p0, const float* dA, const float* dB, const float* dC, int cap,
                            const long* idx, int nidx, double* out) {
    Atoms S{model, na, atoms, cpar, cstride, {}, {}, {npbc_a, npbc_b}, {}, {}};
    for (int k = 0; k < 4; ++k) { S.Qs[k] = Qs[k]; S.QZs[k] = QZs[k]; }
    for (int k = 0; k < 3; ++k) { S.lvA[k] = lvA[k]; S.lvB[k] = lvB[k]; }
    const long nx = n[0], ny = n[1];
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < nidx; ++i) {
        long v = idx[i];
        long ia = v % nx, ib = (v / nx) % ny, ic = v / (nx * ny);
        V3<double> p = v3<double>((double)p0[0] + (double)dA[0] * ia + (double)dB[0] * ib + (double)dC[0] * ic,
                                  (double)p0[1] + (double)dA[1] * ia + (double)dB[1] * ib + (double)dC[1] * ic,
                                  (double)p0[2] + (double)dA[2] * ia + (double)dB[2] * ib + (double)dC[2] * ic);
        double acc[4];
        sample_fe<double>(S, p, acc);
        if (cap) {
            double renorm = 100.0 / std::fabs(acc[3]);
            if (renorm < 1.0) for (int q = 0; q < 4; ++q) acc[q] *= renorm;
        }
        for (int q = 0; q < 4; ++q) out[4 * (size_t)i + q] = acc[q];
    }
}

static Grid make_grid(const float* grid, const int* n, const float* dinvA, const float* dinvB, const float* dinvC) {
    Grid g;
    g.data = grid;
    for (int k = 0; k < 3; ++k) g.n[k] = n[k];
    for (int q = 0; q < 4; ++q) { g.dinv[0][q] = dinvA[q]; g.dinv[1][q] = dinvB[q]; g.dinv[2][q] = dinvC[q]; }
    return g;
}

// Trilinear sample at arbitrary world points (double arithmetic).
void oracle_interp_f64(const float* grid, const int* n, const float* dinvA, const float* dinvB, const float* dinvC,
                       const double* pos, int np, double* out) {
    Grid g = make_grid(grid, n, dinvA, dinvB, dinvC);
    for (int i = 0; i < np; ++i) {
        FE<double> fe = interp<double>(g, v3<double>(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
        out[4 * i] = fe.x; out[4 * i + 1] = fe.y; out[4 * i + 2] = fe.z; out[4 * i + 3] = fe.w;
    }
}

#define RELAX_TILTED(NAME, R)                                                                                      \
    void NAME(const float* grid, const int* n, const float* dinvA, const float* dinvB, const float* dinvC,          \
              const float* pts, int n_scan, int nz, const float* tipA, const float* tipB, const float* tipC,        \
              const float* stiffness, const float* dpos0, const float* relax, const float* surfFF, int max_iter,    \
              float f2conv, int check_before, float* FEs, float* PPs, int damped, int f32_anchors) {                \
        Grid g = make_grid(grid, n, dinvA, dinvB, dinvC);                                                           \
        relax_tilted<R>(g, pts, n_scan, nz, tipA, tipB, tipC, stiffness, dpos0, relax, surfFF, max_iter, f2conv,    \
                        check_before, FEs, PPs, damped, f32_anchors);                                               \
    }
RELAX_TILTED(oracle_relax_tilted_f64, double)
RELAX_TILTED(oracle_relax_tilted_f32, float)

#define RELAX_STROKES(NAME, R)                                                                                     \
    void NAME(const float* grid, const int* n, const float* dinvA, const float* dinvB, const float* dinvC,          \
              const float* pts, int n_scan, int nz, const float* dTip, const float* stiffness, const float* dpos0,  \
              const float* relax, int max_iter, float f2conv, float* FEs, float* disps, float* PPs, int damped) {   \
        Grid g = make_grid(grid, n, dinvA, dinvB, dinvC);                                                           \
        relax_strokes<R>(g, pts, n_scan, nz, dTip, stiffness, dpos0, relax, max_iter, f2conv, FEs, disps, PPs,      \
                         damped);                                                                                   \
    }
RELAX_STROKES(oracle_relax_strokes_f64, double)
RELAX_STROKES(oracle_relax_strokes_f32, float)

#define RELAX_2D(NAME, R)                                                                                          \
    void NAME(const float* grid, const int* n, const float* dinvA, const float* dinvB, const float* dinvC,          \
              const float* pts, int n_scan, int nz, float k_lat, float dh, float dt, float damp, int max_iter,      \
              float f2conv, float* FEs, float* disps, float* PPs) {                                                 \
        Grid g = make_grid(grid, n, dinvA, dinvB, dinvC);                                                           \
        relax_2d<R>(g, pts, n_scan, nz, k_lat, dh, dt, damp, max_iter, f2conv, FEs, disps, PPs);                    \
    }
RELAX_2D(oracle_relax_2d_f64, double)
RELAX_2D(oracle_relax_2d_f32, float)

#define FE_STROKES(NAME, R)                                                                                        \
    void NAME(const float* grid, const int* n, const float* dinvA, const float* dinvB, const float* dinvC,          \
              const float* pts, int n_scan, int nz, const float* tipA, const float* tipB, const float* tipC,        \
              const float* dTip, const float* dpos0, float* FEs) {                                                  \
        Grid g = make_grid(grid, n, dinvA, dinvB, dinvC);                                                           \
        fe_in_strokes_tilted<R>(g, pts, n_scan, nz, tipA, tipB, tipC, dTip, dpos0, FEs);                            \
    }
FE_STROKES(oracle_fe_in_strokes_tilted_f64, double)
FE_STROKES(oracle_fe_in_strokes_tilted_f32, float)

// Fully double-precision variants (double inputs, double grid, double outputs): the
// reference forward in FP64 end to end, used for finite-difference gradients.
void oracle_grid_project_f64g(int model, const double* atoms, const double* cpar, int cstride, int na,
                              const double* Qs, const double* QZs, int npbc_a, int npbc_b, const double* lvA,
                              const double* lvB, const int* n, const double* p0, const double* dA, const double* dB,
                              const double* dC, int cap, double* out) {
    AtomsT<double> S{model, na, atoms, cpar, cstride, {}, {}, {npbc_a, npbc_b}, {}, {}};
    for (int k = 0; k < 4; ++k) { S.Qs[k] = Qs[k]; S.QZs[k] = QZs[k]; }
    for (int k = 0; k < 3; ++k) { S.lvA[k] = lvA[k]; S.lvB[k] = lvB[k]; }
    const long nx = n[0], ny = n[1], nzz = n[2];
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (long ic = 0; ic < nzz; ++ic)
        for (long ib = 0; ib < ny; ++ib)
            for (long ia = 0; ia < nx; ++ia) {
                V3<double> p = v3<double>(p0[0] + dA[0] * ia + dB[0] * ib + dC[0] * ic,
                                          p0[1] + dA[1] * ia + dB[1] * ib + dC[1] * ic,
                                          p0[2] + dA[2] * ia + dB[2] * ib + dC[2] * ic);
                double acc[4];
                sample_fe<double>(S, p, acc);
                if (cap) {
                    double renorm = 100.0 / std::fabs(acc[3]);
                    if (renorm < 1.0) for (int q = 0; q < 4; ++q) acc[q] *= renorm;
                }
                double* o = out + 4 * ((ic * ny + ib) * nx + ia);
                for (int q = 0; q < 4; ++q) o[q] = acc[q];
            }
}

void oracle_relax_tilted_f64g(const double* grid, const int* n, const double* dinvA, const double* dinvB,
                              const double* dinvC, const double* pts, int n_scan, int nz, const double* tipA,
                              const double* tipB, const double* tipC, const double* stiffness, const double* dpos0,
                              const double* relax, const double* surfFF, int max_iter, double f2conv,
                              int check_before, double* FEs, double* PPs, int f32_anchors) {
    GridT<double> g;
    g.data = grid;
    for (int k = 0; k < 3; ++k) g.n[k] = n[k];
    for (int q = 0; q < 4; ++q) { g.dinv[0][q] = dinvA[q]; g.dinv[1][q] = dinvB[q]; g.dinv[2][q] = dinvC[q]; }
    relax_tilted<double>(g, pts, n_scan, nz, tipA, tipB, tipC, stiffness, dpos0, relax, surfFF, max_iter, f2conv,
                         check_before, FEs, PPs, 0, f32_anchors);
}

// relaxStrokesTiltedMorseDirect  (kernels/AFM.cl:1419-1511); cMs rows (R0,E0,K,0).
#define RELAX_DIRECT(NAME, R)                                                                                      \
    void NAME(const float* atoms, const float* cMs, int na, const float* Qs, const float* QZs, const float* pts,    \
              int n_scan, int nz, const float* tipA, const float* tipB, const float* tipC, const float* stiffness,  \
              const float* dpos0, const float* relax, const float* surfFF, int max_iter, float f2conv, float* FEs,  \
              float* PPs) {                                                                                         \
        relax_direct<R>(atoms, cMs, na, Qs, QZs, pts, n_scan, nz, tipA, tipB, tipC, stiffness, dpos0, relax, surfFF, \
                        max_iter, f2conv, FEs, PPs);                                                                \
    }
RELAX_DIRECT(oracle_relax_direct_f64, double)
RELAX_DIRECT(oracle_relax_direct_f32, float)

// The same in double end to end (double inputs and outputs), for finite differences.
void oracle_relax_direct_f64d(const double* atoms, const double* cMs, int na, const double* Qs, const double* QZs,
                              const double* pts, int n_scan, int nz, const double* tipA, const double* tipB,
                              const double* tipC, const double* stiffness, const double* dpos0, const double* relax,
                              const double* surfFF, int max_iter, double f2conv, double* FEs, double* PPs) {
    relax_direct<double>(atoms, cMs, na, Qs, QZs, pts, n_scan, nz, tipA, tipB, tipC, stiffness, dpos0, relax, surfFF,
                         max_iter, f2conv, FEs, PPs);
}

// Implicit-equilibrium VJP of the direct path in double (kernels/AFM.cl:1711-2029 semantics):
// per state, gates in the reference order — non-finite PP/u (2), PP.w <= 0 (1),
// residual |G| >= 5e-4 (1), non-finite Hessian (2), singular J (4), non-finite lambda (2),
// smallest eigenvalue of -J <= 1e-5 (3), condition number >= 1e6 (4); J = -H + R^T (dg/dd) R,
// b = -H R^T u - F u.w, J lambda = b; w = R^T u - lambda; per atom
// dL/da = H_i w + F_i u.w, dL/dR0 = w.dF/dR0 + u.w dE/dR0, ... (theta order x,y,z,R0,E0,Q).
// A gated state contributes nothing.  diag[s] = (residual, lambda_min, cond, status).
void oracle_direct_vjp_f64(const double* atoms, const double* cMs, int na, const double* Qs, const double* QZs,
                           const double* pts, int n_scan, int nz, const double* tipA, const double* tipB,
                           const double* tipC, const double* stiffness, const double* dpos0, const double* surfFF,
                           const double* PPs, const double* U, double* lambdas, double* diag, double* grad) {
    double Qk[4], QZk[4];
    for (int k = 0; k < 4; ++k) { Qk[k] = Qs[k] * kCOULOMB; QZk[k] = QZs[k]; }
    const double Rm[3][3] = {{tipA[0], tipA[1], tipA[2]}, {tipB[0], tipB[1], tipB[2]}, {tipC[0], tipC[1], tipC[2]}};
    for (int i = 0; i < 6 * na; ++i) grad[i] = 0.0;
    std::vector<PairD> pr(na);
    for (int ip = 0; ip < n_scan; ++ip)
        for (int iz = 0; iz < nz; ++iz) {
            const size_t st = (size_t)ip * nz + iz;
            const double* pp = PPs + 4 * st;
            const double* u = U + 4 * st;
            double* lam = lambdas + 4 * st;
            double* dg = diag + 4 * st;
            lam[0] = lam[1] = lam[2] = lam[3] = 0.0;
            dg[0] = dg[1] = dg[2] = 0.0;
            bool fin = true;
            for (int q = 0; q < 4; ++q) fin = fin && std::isfinite(pp[q]) && std::isfinite(u[q]);
            if (!fin) { dg[3] = 2; continue; }
            if (!(pp[3] > 0.0)) { dg[3] = 1; continue; }
            const double q[3] = {pp[0], pp[1], pp[2]};
            double tip[3];
            for (int k = 0; k < 3; ++k) tip[k] = pts[4 * (size_t)ip + k] + tipC[k] * tipC[3] * (double)iz;
            double F[3] = {0, 0, 0}, H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            for (int j = 0; j < na; ++j) {
                pair_direct(q, atoms + 4 * (size_t)j, cMs + 4 * (size_t)j, Qk, QZk, pr[j]);
                for (int a = 0; a < 3; ++a) {
                    F[a] += pr[j].F[a];
                    for (int b = 0; b < 3; ++b) H[a][b] += pr[j].H[a][b];
                }
            }
            // total force G and tip-spring Jacobian in tip coordinates
            double dw[3] = {q[0] - tip[0], q[1] - tip[1], q[2] - tip[2]}, d[3];
            for (int a = 0; a < 3; ++a) d[a] = Rm[a][0] * dw[0] + Rm[a][1] * dw[1] + Rm[a][2] * dw[2];
            double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            r = std::max(r, 1e-10);
            const double phi = stiffness[3] * (1.0 - dpos0[3] / r);
            const double beta = stiffness[3] * dpos0[3] / (r * r * r);
            double gt[3], Jt[3][3];
            for (int a = 0; a < 3; ++a) {
                gt[a] = (d[a] - dpos0[a]) * stiffness[a] + d[a] * (stiffness[3] * (r - dpos0[3]) / r);
                for (int b = 0; b < 3; ++b) Jt[a][b] = beta * d[a] * d[b] + (a == b ? stiffness[a] + phi : 0.0);
            }
            double G[3], J[3][3];
            for (int a = 0; a < 3; ++a) {
                G[a] = F[a] + Rm[0][a] * gt[0] + Rm[1][a] * gt[1] + Rm[2][a] * gt[2] + tipC[a] * surfFF[0];
                for (int b = 0; b < 3; ++b) {
                    double acc = -H[a][b];
                    for (int m = 0; m < 3; ++m)
                        for (int n = 0; n < 3; ++n) acc += Rm[m][a] * Jt[m][n] * Rm[n][b];
                    J[a][b] = acc;
                }
            }
            const double res = std::sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
            dg[0] = res;
            if (!(res < 5e-4)) { dg[3] = 1; continue; }
            bool hfin = true;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) hfin = hfin && std::isfinite(H[a][b]);
            if (!hfin) { dg[3] = 2; continue; }
            double uw[3];
            for (int a = 0; a < 3; ++a) uw[a] = Rm[0][a] * u[0] + Rm[1][a] * u[1] + Rm[2][a] * u[2];
            double bv[3];
            for (int a = 0; a < 3; ++a) bv[a] = -(H[a][0] * uw[0] + H[a][1] * uw[1] + H[a][2] * uw[2]) - F[a] * u[3];
            // cofactors and determinant of J
            double cof[3][3];
            cof[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
            cof[0][1] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
            cof[0][2] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
            cof[1][0] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
            cof[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
            cof[1][2] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
            cof[2][0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
            cof[2][1] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
            cof[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            const double det = J[0][0] * cof[0][0] + J[0][1] * cof[0][1] + J[0][2] * cof[0][2];
            const double scale = std::max(std::max(std::fabs(J[0][0]), std::fabs(J[1][1])),
                                          std::max(std::fabs(J[2][2]), 1e-20));
            if (std::fabs(det) < 1e-18 * scale * scale * scale) { dg[3] = 4; continue; }
            double l[3];
            for (int a = 0; a < 3; ++a) l[a] = (cof[0][a] * bv[0] + cof[1][a] * bv[1] + cof[2][a] * bv[2]) / det;
            if (!(std::isfinite(l[0]) && std::isfinite(l[1]) && std::isfinite(l[2]))) { dg[3] = 2; continue; }
            double S[3][3];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) S[a][b] = -0.5 * (J[a][b] + J[b][a]);
            const double lmin = eigmin_sym3(S);
            dg[1] = lmin;
            if (!(lmin > 1e-5)) { dg[3] = 3; continue; }
            double jn = 0.0, an = 0.0;
            for (int a = 0; a < 3; ++a) {
                jn = std::max(jn, std::fabs(J[a][0]) + std::fabs(J[a][1]) + std::fabs(J[a][2]));
                an = std::max(an, std::fabs(cof[0][a]) + std::fabs(cof[1][a]) + std::fabs(cof[2][a]));
            }
            const double cond = jn * an / std::max(std::fabs(det), 1e-300);
            dg[2] = cond;
            if (!(cond < 1e6)) { dg[3] = 4; continue; }
            dg[3] = 0;
            lam[0] = l[0]; lam[1] = l[1]; lam[2] = l[2];
            const double w[3] = {uw[0] - l[0], uw[1] - l[1], uw[2] - l[2]};
            for (int j = 0; j < na; ++j) {
                const PairD& P = pr[j];
                double* g = grad + 6 * (size_t)j;
                for (int a = 0; a < 3; ++a)
                    g[a] += P.H[a][0] * w[0] + P.H[a][1] * w[1] + P.H[a][2] * w[2] + P.F[a] * u[3];
                g[3] += w[0] * P.dF_dR0[0] + w[1] * P.dF_dR0[1] + w[2] * P.dF_dR0[2] + u[3] * P.dE_dR0;
                g[4] += w[0] * P.dF_dE0[0] + w[1] * P.dF_dE0[1] + w[2] * P.dF_dE0[2] + u[3] * P.dE_dE0;
                g[5] += w[0] * P.dF_dQ[0] + w[1] * P.dF_dQ[1] + w[2] * P.dF_dQ[2] + u[3] * P.dE_dQ;
            }
        }
}

}  // extern "C"
