/*
 * ppm_api.h — C ABI of the MI355X (gfx950) probe-particle AFM library
 *             (libppm_hip.so, built from spammm_amd/csrc/).
 *
 * Drop-in boundary for the probe-particle-model AFM path of SPAMMM.  Every
 * entry point below replaces one OpenCL kernel launch (or host NumPy stage) of
 * the reference; the replaced interface is cited on each declaration.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer (HIP global memory) owned by the
 *     caller; the library allocates nothing on these paths and keeps no state
 *     between calls apart from compiled code.
 *   - `stream` is a hipStream_t (NULL = the default stream).  Calls are
 *     stream-ordered and return after the launch; they never synchronize the
 *     device.  Calls on different streams may run concurrently.
 *   - Return value: PPM_OK (0) or a PPM_ERR_* code; ppm_last_error() returns a
 *     thread-local message describing the last failure on the calling thread.
 *   - float4 arrays are packed as 4 consecutive floats (x,y,z,w), 16-byte aligned
 *     (the relaxation entry points reject misaligned points and outputs).
 *   - Grid layout: (nz, ny, nx, 4) float, x fastest — the texel order of the
 *     reference's RGBA32F image3d (kernels/AFM.cl:989, write_imagef at (ia,ib,ic)).
 *   - Scan layout: point index k = ix*ny + iy (ix-major, spammm/SPM/AFM.py:977-980);
 *     per-(point, z) outputs at index k*nz + iz, iz=0 = first (highest) tip position.
 */
#ifndef PPM_API_H
#define PPM_API_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PPM_API_VERSION 2

/* status codes */
#define PPM_OK                 0
#define PPM_ERR_INVALID_ARG    1
#define PPM_ERR_HIP            2
#define PPM_ERR_UNSUPPORTED    3

/* pair models */
#define PPM_MODEL_LJ     0   /* cpar row = (c6, c12)            getLJ     kernels/AFM.cl:738-745 */
#define PPM_MODEL_MORSE  1   /* cpar row = (R0, E0, K<0, unused) getMorse kernels/AFM.cl:747-757 */

/* Grid geometry.  Voxel (ia,ib,ic) sits at p0 + dA*ia + dB*ib + dC*ic
 * (kernels/AFM.cl:959).  The normalized sampling coordinate of a world point pos
 * is s_a = dot((pos,1), dinvA) etc. (interpFE, kernels/AFM.cl:100-105); the .w
 * entries carry -origin/L + 0.5/n (spammm/SPM/AFM.py:275-294). */
typedef struct PpmGridSpec {
    int   nx, ny, nz, reserved;
    float p0[4], dA[4], dB[4], dC[4];
    float dinvA[4], dinvB[4], dinvC[4];
} PpmGridSpec;

/* Tip / probe parameter record (AFMulator.DEFAULT_*, spammm/SPM/AFM.py:296-304). */
typedef struct PpmTip {
    float tipA[4], tipB[4], tipC[4];  /* tip frame rows; tipC.w = dtip (z step per slice) */
    float stiffness[4];               /* (k_x, k_y, k_z, k_rad), negative = restoring [eV/A^2] */
    float dpos0[4];                   /* probe rest offset (x,y,z) in tip frame, .w = bond length L */
    float relax[4];                   /* (dt, damp, unused, unused) */
    float surfFF[4];                  /* .x = constant force along tipC */
} PpmTip;

/* Relaxation control (compile-time N_RELAX_STEP_MAX / F2CONV of kernels/AFM.cl:121-122
 * made runtime). */
typedef struct PpmRelaxCtl {
    int   max_iter;       /* FIRE iterations per z slice (reference: 128) */
    float f2conv;         /* stop when |F|^2 < f2conv (reference: 1e-8) */
    int   check_before;   /* 0: test after the update (relaxStrokesTilted, kernels/AFM.cl:544);
                             1: test before it (relaxStrokesTiltedMorseDirect, kernels/AFM.cl:1478) */
    int   scan_ny;        /* >0: points form an ix-major raster with this many iy per ix
                             (enables 8x8 pixel tiles per wavefront); 0: no raster */
    int   f32_anchors;    /* FP64 relaxation only: 1 = tip anchors advance as the FP32 kernels
                             advance them (tip += dTip in float32, kernels/AFM.cl:556), so the
                             FP64 scan follows the FP32 scan's exact heights; 0 = in FP64 */
    int   damped;         /* integrator of the relaxation loops: 0 = FIRE (update_FIRE, the
                             reference's default OPT_FIRE=1 build); 1 = damped velocity
                             v *= (1 - damp) with constant dt, damp (its OPT_FIRE=0 build,
                             AFMulator(use_fire=False), kernels/AFM.cl:378-382, :536-540) */
    int   lanes;          /* lanes per scan point of ppm_relax_strokes_tilted*, ppm_relax_strokes and
                             ppm_relax_strokes_2d:
                             1, 2 or 4, or 0 = the library's choice (the fastest for a launch that has
                             the GPU to itself: about two waves per SIMD).  Results do not depend on
                             it; 1 issues the fewest instructions, for a relaxation that shares the
                             GPU with other work (consecutive scans on two streams). */
} PpmRelaxCtl;

/* One molecule of a batch (an array of these lives in device memory).  The molecule's grid
 * geometry is that of PpmGridSpec; only the z planes [klo, klo + nzw) are computed and stored
 * (at grid_off float4s into the batch grid buffer, plane-major, x fastest).  Its scan is the
 * ix-major raster start(ix,iy) = (scan_p0 + scan_da*ix) + scan_db*iy, ix < scan_nx,
 * iy < scan_ny; per-pixel outputs start at pixel pix_off.  Work queue items
 * [item_off, item_off + n_grid_items) are the molecule's grid tiles and
 * [relax_item_off, relax_item_off + n_relax_items) its relaxation blocks (relax_item_off = 0:
 * right after the tiles).  Across molecules both kinds of segment start in increasing order of the
 * molecule index and every relaxation segment comes after its own molecule's tiles; the segments
 * of different molecules may interleave (e.g. tiles of m + k before relaxation blocks of m). */
typedef struct PpmMolDesc {
    int atom_off, na;
    int nx, ny, nz;
    int klo, nzw;
    int npbc_a, npbc_b;
    int pix_off, scan_nx, scan_ny;
    int item_off, n_grid_items, n_relax_items, relax_item_off;
    long long grid_off;
    float p0[4], dA[4], dB[4], dC[4];
    float dinvA[4], dinvB[4], dinvC[4];
    float lvA[4], lvB[4];
    float scan_p0[4], scan_da[4], scan_db[4];
} PpmMolDesc;

/* Settings shared by every molecule of a batch. */
typedef struct PpmBatchParams {
    int model;          /* PPM_MODEL_LJ or PPM_MODEL_MORSE (cpar rows of 2 or 4 floats) */
    int cap;            /* |E|>100 rescale of the grid */
    int nz;             /* z-steps of every scan */
    int nmol, nitems;
    int vz;             /* grid voxels per thread along z: 4 or 8 */
    int tile_w;         /* pixels along iy of a wavefront's pixel tile: 4, 8 or 16 (64/tile_w along ix) */
    int mode;           /* PPM_BATCH_FUSED, PPM_BATCH_GRID_ONLY or PPM_BATCH_RELAX_ONLY */
    int df_taps, df_nz_out;
    int reserved[2];
    float Qs[4], QZs[4];
    PpmTip tip;
    PpmRelaxCtl ctl;
} PpmBatchParams;

#define PPM_BATCH_FUSED       0   /* grid tiles and relaxation blocks from one work queue */
#define PPM_BATCH_GRID_ONLY   1
#define PPM_BATCH_RELAX_ONLY  2

/* Library version (PPM_API_VERSION). */
int ppm_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* ppm_last_error(void);

/* A non-blocking HIP stream on the current device (*stream), and its release.  No counterpart in
 * the reference (it runs one in-order OpenCL queue per AFMulator); for callers that keep several
 * scans in flight at once.  A stream created here takes the least used hardware queue of the
 * process when it is created, whereas streams handed out from a pool that was created in one go
 * can share a hardware queue and then run one after another.  ppm_stream_destroy may be called
 * with work still queued on the stream; its resources are released when that work is done. */
int ppm_stream_create(void** stream);
int ppm_stream_destroy(void* stream);

/* Force-field grid projection.
 * Replaces evalLJC_QZs_toImg (kernels/AFM.cl:916-991) and evalMorseC_QZs_toImg
 * (kernels/AFM.cl:996-1056) as launched by AFMulator.make_forcefield
 * (spammm/SPM/AFM.py:897-946), extended with periodic x,y images in the
 * convention of make_MorseFF (kernels/gridFF.cl:938-996).
 *   atoms    (na,4)  x,y,z,q                     [A, e]
 *   cpar     (na,2) for PPM_MODEL_LJ, (na,4) for PPM_MODEL_MORSE
 *   Qs, QZs  host float[4]: tip charges [e] and their z offsets [A]
 *            (COULOMB_CONST is applied inside, kernels/AFM.cl:961)
 *   npbc_a, npbc_b  image range [-n,n] along lvA, lvB (host float[3]); 0 = no images
 *   cap      nonzero: rescale all four channels by 100/|E| when |E|>100
 *            (kernels/AFM.cl:984-985)
 *   grid     (nz,ny,nx,4) output. */
int ppm_grid_project(int model, const float* atoms, const float* cpar, int na,
                     const PpmGridSpec* spec, const float Qs[4], const float QZs[4],
                     int npbc_a, int npbc_b, const float lvA[3], const float lvB[3],
                     int cap, float* grid, void* stream);

/* ppm_grid_project with each voxel's pair terms, evaluated in FP32 exactly as there, summed in
 * FP64 and rounded once to the FP32 output.  Same arguments, pair laws, images and rescale
 * (evalLJC_QZs_toImg / evalMorseC_QZs_toImg, kernels/AFM.cl:916-1056, whose FP32 sum it
 * refines).  Above a neutral molecule the thousands of terms of a voxel cancel to ~1e-2 of
 * their magnitude and the FP32 running sum dominates the field's error; this form removes that
 * part (3-6x lower error there) at ~1.55x the time (8 voxels per thread), for uses bound by
 * the field's own accuracy. */
int ppm_grid_project_acc64(int model, const float* atoms, const float* cpar, int na,
                           const PpmGridSpec* spec, const float Qs[4], const float QZs[4],
                           int npbc_a, int npbc_b, const float lvA[3], const float lvB[3],
                           int cap, float* grid, void* stream);

/* ppm_grid_project of the z planes [klo, klo + nzw) only: grid (nzw,ny,nx,4) holds the voxels
 * (ia, ib, klo + k) of the spec's grid, bit for bit the values ppm_grid_project stores there
 * (same pair laws, images and rescale, evalLJC_QZs_toImg / evalMorseC_QZs_toImg,
 * kernels/AFM.cl:916-1056).  No reference counterpart: the reference always projects the whole
 * image (AFMulator.make_forcefield, spammm/SPM/AFM.py:897-946); a scan that descends over a few
 * A of a tall cell samples only the planes around its probe heights, and
 * ppm_relax_strokes_tilted_window relaxes on such a window.  0 <= klo, klo + nzw <= nz. */
int ppm_grid_project_window(int model, const float* atoms, const float* cpar, int na,
                            const PpmGridSpec* spec, int klo, int nzw, const float Qs[4], const float QZs[4],
                            int npbc_a, int npbc_b, const float lvA[3], const float lvB[3],
                            int cap, float* grid, void* stream);

/* ppm_grid_project over a base: grid = cap(base + pair sum of `atoms`), the |E|>100 rescale
 * (if cap) applied to the total.  base: a raw (cap = 0) grid of other atoms on the same spec,
 * e.g. a substrate computed once and shared by many molecule placements; it must not be the
 * same buffer as grid.  No reference counterpart (the reference projects one atom list per
 * image, kernels/AFM.cl:916-1056); with base = a grid of atoms S, the result equals
 * ppm_grid_project of S + atoms up to FP32 summation order. */
int ppm_grid_project_add(int model, const float* atoms, const float* cpar, int na,
                         const PpmGridSpec* spec, const float Qs[4], const float QZs[4],
                         int npbc_a, int npbc_b, const float lvA[3], const float lvB[3], int cap,
                         const float* base, float* grid, void* stream);

/* ppm_grid_project in FP64: double atoms (na,4), cpar ((na,2) LJ c6,c12 / (na,4) Morse R0,E0,K,0),
 * tip charges, lattice vectors and output grid (nz,ny,nx,4); voxel positions from the spec in
 * double.  Same pair laws, images and rescale as ppm_grid_project (evalLJC_QZs_toImg /
 * evalMorseC_QZs_toImg, kernels/AFM.cl:916-1056, images as kernels/gridFF.cl:938-996) at the
 * precision of the reference's CPU64 test oracle (tests/SPM/test_afm_morse.py:410-624); for
 * finite-difference gradient checks and reference-quality fields, not a hot path. */
int ppm_grid_project_f64(int model, const double* atoms, const double* cpar, int na,
                         const PpmGridSpec* spec, const double Qs[4], const double QZs[4],
                         int npbc_a, int npbc_b, const double lvA[3], const double lvB[3],
                         int cap, double* grid, void* stream);

/* ppm_relax_strokes_tilted in FP64: double grid (nz,ny,nx,4) (e.g. from ppm_grid_project_f64),
 * double points (n_scan,4), double FEs / PPs (n_scan,nz,4) (PPs may be NULL); same semantics
 * as ppm_relax_strokes_tilted (relaxStrokesTilted, kernels/AFM.cl:478-562) with every operation
 * in FP64, at the precision of the reference's CPU64 test oracle
 * (tests/SPM/test_afm_morse.py:410-624).  Not a hot path: the reference-precision forward for
 * finite-difference gradient checks. */
int ppm_relax_strokes_tilted_f64(const double* grid, const PpmGridSpec* spec,
                                 const double* points, int n_scan, int nz,
                                 const PpmTip* tip, const PpmRelaxCtl* ctl,
                                 double* FEs, double* PPs, void* stream);

/* ---- FDBM force-field producer (spammm/SPM/AFM.py:3304-3566, :4204-4253; kernels/AFM.cl:1151-1303) ----
 * Scalar fields are (nx,ny,nz) C-order float32 (z fastest, the reference's host layout);
 * grid spacing `step` is isotropic.  FFTs are FP32 hipFFT real<->complex transforms. */
enum {
    PPM_FFT_CORRELATE = 0,          /* out = scale * IFFT(FFT(a) . conj(FFT(b)))            */
    PPM_FFT_CONVOLVE = 1,           /* out = scale * IFFT(FFT(a) . FFT(b))                  */
    PPM_FFT_POISSON = 2,            /* out = scale * IFFT(FFT(a) / k^2), k = 0 -> 0; b unused */
    PPM_FFT_POISSON_CONVOLVE = 3    /* out = scale * IFFT(FFT(a) . FFT(b) / k^2), k = 0 -> 0 */
};

/* Bytes of device workspace of ppm_fdbm_fft_combine: two half spectra (nx*ny*(nz/2+1) complex). */
size_t ppm_fdbm_workspace_size(int nx, int ny, int nz);

/* The FFT products of the FDBM stage (IFFT normalised by 1/(nx ny nz) as NumPy's ifftn):
 * Pauli overlap = CORRELATE(rho, tip_tot) with scale dV (compute_pauli_overlap,
 * spammm/SPM/AFM.py:4092-4103, corr_conj :3469-3478); Poisson potential = POISSON(rho) with
 * scale 4 pi COULOMB_CONST (fft_poisson :3580-3587, poisson :3458-3467); fused electrostatics =
 * POISSON_CONVOLVE(rho_diff, flip(tip_del)) with scale 4 pi COULOMB_CONST dV
 * (es_fused_from_rho_cl :3543-3565, fdbm_mul_poisson_tip_c64 kernels/AFM.cl:1287-1303);
 * electrostatic convolution = CONVOLVE(V, flip(tip_del)) with scale dV
 * (compute_es_conv_field :4153-4175).  a, b, out: device (nx,ny,nz); out may not alias a or b.
 * hipFFT plans are created on the first call per shape and device and kept until
 * ppm_fdbm_release_plans (plan creation allocates hipFFT's own work area); a call's stream
 * binding and transforms run under one lock, so calls from several threads on different
 * streams each run on their own stream. */
int ppm_fdbm_fft_combine(int mode, int nx, int ny, int nz, float step, const float* a, const float* b,
                         float scale, float* out, void* workspace, size_t workspace_bytes, void* stream);

/* Destroys every cached hipFFT plan of ppm_fdbm_fft_combine (on every device; the next call
 * creates them again).  Call it with no FFT work of the library in flight, e.g. at shutdown. */
int ppm_fdbm_release_plans(void);

/* field[i] = A * max(field[i], clip)^beta over n values (fdbm_scale_pauli_pow_f32,
 * kernels/AFM.cl:1258-1267; scale_pauli_field spammm/SPM/AFM.py:4106-4128). */
int ppm_fdbm_pauli_scale(float* field, long long n, float A, float beta, float clip, void* stream);

/* dst (shape) = roll(pad(src (src_shape) at offset), -roll): dst[i] = padded[(i + roll) mod n]
 * (fdbm_pad_roll_f32, kernels/AFM.cl:1202-1221). */
int ppm_fdbm_pad_roll(const float* src, const int src_shape[3], float* dst, const int shape[3],
                      const int offset[3], const int roll[3], void* stream);

/* dst[ix,iy,iz] = src[nx-1-ix, ny-1-iy, nz-1-iz] (fdbm_flip3_f32, kernels/AFM.cl:1223-1233). */
int ppm_fdbm_flip3(const float* src, float* dst, const int shape[3], void* stream);

/* London dispersion on the grid p = origin + step*(ix,iy,iz): out (+)= -sum_a w_a / (|p - r_a|^2 + RA^2)^3
 * with atoms (na,4) = (x,y,z, w = sqrt(C6_a C6_tip)) (evalDispersion_toImg + getLondon,
 * kernels/AFM.cl:1066-1113, :804-814; compute_dispersion_to_img_cl spammm/SPM/AFM.py:431-462). */
int ppm_fdbm_dispersion(const float* atoms, int na, const float origin[3], float step, const int shape[3],
                        float RA, int accumulate, float* out, void* stream);

/* E = e1 + e2 + e3 (any may be NULL); grid (nz,ny,nx,4) = (-dE/dx, -dE/dy, -dE/dz, E) by periodic
 * central differences (fdbm_compose_E_to_img + gradient_central_diff, kernels/AFM.cl:1269-1285,
 * :1151-1196; compose_E_and_gradient_fast_cl spammm/SPM/AFM.py:464-483). */
int ppm_fdbm_compose_gradient(const float* e1, const float* e2, const float* e3, const int shape[3],
                              float step, float* grid, void* stream);

/* ---- 2D Ewald electrostatics of periodic slabs (spammm/surfaces/Ewald2D.py,
 * spammm/surfaces/SurfaceEwald.py, kernels/surface.cl:1036-1381) ----
 *   phi(rho,z) = C [ -(2 pi/A) sum_i q_i |z - z_i|
 *                    + sum_{G!=0} sum_i (2 pi/(A|G|)) q_i Re e^{iG.(rho-rho_i)} e^{-|G||z-z_i|} ],
 * G = h b1 + k b2 with |h|,|k| <= n_harm (make_reciprocal_2d / generate_G_vectors,
 * Ewald2D.py:39-84), C = 14.3996448915 eV A / e^2.  Lattice vectors a, b are host double[2].
 * Ions are (n,4) float (x,y,z,q) in device memory, grouped by layer: layer l holds
 * ions[layer_off[l] .. layer_off[l+1]) and all of them share one z (layer_off: device int[n_layers+1]). */
enum {
    PPM_EWALD_FULL = 0,      /* eval_potential_full (kernels/surface.cl:1172-1246), any z             */
    PPM_EWALD_VACUUM = 1     /* eval_potential_vacuum (kernels/surface.cl:1103-1157), z above the ions */
};

/* Number of G vectors: (2 n_harm + 1)^2 - 1. */
int ppm_ewald_num_g(int n_harm);

/* Bytes of the prepared-system buffer of ppm_ewald_prepare (1 <= n_layers, 1 <= n_harm <= 64). */
size_t ppm_ewald_system_bytes(int n_layers, int n_harm);

/* Byte offset, inside the prepared-system buffer, of the int count of kept G (written by
 * ppm_ewald_prepare; read it with a device-to-host copy).  Each kept G stands for the pair +-G,
 * so the count is half the number of G vectors summed. */
size_t ppm_ewald_kept_count_offset(int n_layers, int n_harm);

/* Prepares a slab (SurfaceEwaldCL.prepare_system, SurfaceEwald.py:165-234, whose
 * compute_ewald_coefficients kernel is kernels/surface.cl:1036-1086): per-layer structure factors
 * W_l(G) = (2 pi/(A|G|)) sum_{i in l} q_i e^{-iG.rho_i} in FP64, layer heights and charges, and the
 * ordered list of G kept for evaluation: those with max_l |W_l(G)| > drop_tol * max |W| (drop_tol 0
 * keeps every G with a nonzero structure factor), one per pair +-G (h > 0, or h = 0 and k > 0:
 * real charges give W_l(-G) = conj W_l(G), so the pair's two terms are equal and are evaluated
 * once, doubled).  system: device, ppm_ewald_system_bytes. */
int ppm_ewald_prepare(const float* ions, const int* layer_off, int n_layers, const double a[2],
                      const double b[2], int n_harm, double drop_tol, void* system, size_t system_bytes,
                      void* stream);

/* phi (np floats, eV/e) at points (np,4) = (x,y,z,-) (SurfaceEwaldCL.eval_full / eval_vacuum,
 * SurfaceEwald.py:236-347), FP64 accumulation; mode PPM_EWALD_FULL or PPM_EWALD_VACUUM.  system:
 * prepared by ppm_ewald_prepare with the same n_layers, n_harm and lattice (system_bytes is checked
 * against ppm_ewald_system_bytes, so mismatched sizes are rejected instead of read past). */
int ppm_ewald_eval(int mode, const float* points, int np, const void* system, size_t system_bytes, int n_layers,
                   int n_harm, const double a[2], const double b[2], float* phi, void* stream);

/* Direct sum over the images n a + m b: square = 0 takes n^2 + m^2 <= n_rep^2
 * (SurfaceEwaldCL.eval_brute, SurfaceEwald.py:349-405; eval_potential_brute
 * kernels/surface.cl:1260-1305), square = 1 every |n|,|m| <= n_rep (eval_potential_brute,
 * Ewald2D.py:176-204); FP64 accumulation. */
int ppm_ewald_brute(const float* points, int np, const float* ions, int n_ions, const double a[2],
                    const double b[2], int n_rep, int square, float* phi, void* stream);

/* Finite-cluster Coulomb sum without images (SurfaceEwaldCL.eval_cluster, SurfaceEwald.py:407-449;
 * eval_potential_cluster kernels/surface.cl:1321-1381); FP64 accumulation. */
int ppm_ewald_cluster(const float* points, int np, const float* ions, int n_ions, float* phi, void* stream);

/* Bytes of device workspace of ppm_ewald_tip_grid. */
size_t ppm_ewald_grid_workspace_size(int n_harm, int nz);

/* The slab's electrostatics as a probe force-field grid: grid (nz,ny,nx,4) (+)= (F, E) with
 * E = sum_k Qs_k phi(r + QZs_k z) and F = -grad E (analytic), at r = p0 + ia dA + ib dB + ic dC.
 * Requires dA, dB in the xy plane and dC along z.  Same voxel layout and tip-charge convention
 * as ppm_grid_project (the Coulomb sites of evalLJC_QZs_toImg, kernels/AFM.cl:916-991), with the
 * slab's periodic sum done by the Ewald series instead of a finite image sum; no energy cap. */
int ppm_ewald_tip_grid(const void* system, size_t system_bytes, int n_layers, int n_harm, const double a[2],
                       const double b[2], const PpmGridSpec* spec, const float Qs[4], const float QZs[4],
                       int accumulate, void* workspace, size_t workspace_bytes, float* grid, void* stream);

/* Relaxed scan with a tilted tip frame.
 * Replaces relaxStrokesTilted (kernels/AFM.cl:478-562) as launched by
 * AFMulator.run_scan (spammm/SPM/AFM.py:950-1010).
 *   points (n_scan,4) tip start positions; FEs (n_scan,nz,4) tip-rotated sample FE.
 *   PPs    (n_scan,nz,4) or NULL: probe position and +iterations (converged) /
 *          -iterations (not converged or non-finite), the telemetry convention of
 *          relaxStrokesTiltedMorseDirect (kernels/AFM.cl:1494-1505). */
int ppm_relax_strokes_tilted(const float* grid, const PpmGridSpec* spec,
                             const float* points, int n_scan, int nz,
                             const PpmTip* tip, const PpmRelaxCtl* ctl,
                             float* FEs, float* PPs, void* stream);

/* ppm_relax_strokes_tilted with the df operator of ppm_z_stencil applied to each point's Fz
 * column in the same launch (the Fz -> df conversion of compute_df / compute_df_amp / get_df,
 * spammm/SPM/AFM.py:3048-3089, :2591-2599, on the device without a second pass over FEs):
 *   df[k*nz_out + j] = sum_{t<taps} W[j*taps + t] * FEs[k, clamp(start[j] + t, 0, nz-1)].z,
 * accumulated as ppm_z_stencil does (its results bit for bit).  W (nz_out*taps) float and
 * start (nz_out) int: device arrays; df (n_scan, nz_out). */
int ppm_relax_strokes_tilted_df(const float* grid, const PpmGridSpec* spec,
                                const float* points, int n_scan, int nz,
                                const PpmTip* tip, const PpmRelaxCtl* ctl,
                                float* FEs, float* PPs, const float* W, const int* start, int taps,
                                int nz_out, float* df, void* stream);

/* ppm_relax_strokes_tilted on a z window of the grid: `grid` holds only planes [klo, klo + nzw)
 * (as ppm_grid_project_window stores them); sampling coordinates and REPEAT addressing are those
 * of the whole grid `spec`, so wherever every probe samples inside the window the results equal
 * ppm_relax_strokes_tilted on the whole grid bit for bit.  A sample that needs a plane outside
 * the window reads the nearest stored plane and sets *flag (device int, caller-zeroed) nonzero:
 * the results are then not those of the whole grid and the caller recomputes without a window.
 * relaxStrokesTilted semantics (kernels/AFM.cl:478-562); no reference counterpart for the
 * window.  nzw >= 2, 0 <= klo, klo + nzw <= nz. */
int ppm_relax_strokes_tilted_window(const float* grid, const PpmGridSpec* spec, int klo, int nzw,
                                    const float* points, int n_scan, int nz, const PpmTip* tip,
                                    const PpmRelaxCtl* ctl, float* FEs, float* PPs, int* flag, void* stream);

/* Relaxed scan, untilted tip, with displacement output.
 * Replaces relaxStrokes (kernels/AFM.cl:345-394) as launched by
 * AFMulator.scan_fdbm (spammm/SPM/AFM.py:493-642).
 *   dTip, stiffness, dpos0, relax: host float[4].
 *   FEs (n_scan,nz,4): sample taken at the start of the last FIRE iteration;
 *   disps (n_scan,nz,4) or NULL: pos - (tip + dpos0); PPs as above or NULL. */
int ppm_relax_strokes(const float* grid, const PpmGridSpec* spec,
                      const float* points, int n_scan, int nz,
                      const float dTip[4], const float stiffness[4], const float dpos0[4],
                      const float relax[4], const PpmRelaxCtl* ctl,
                      float* FEs, float* disps, float* PPs, void* stream);

/* Lateral-only relaxed scan at fixed heights.
 * Replaces relaxStrokes2D (kernels/AFM.cl:396-450) as launched by AFMulator.scan_fdbm_2d
 * (spammm/SPM/AFM.py:644-701): for point k with anchor (ax, ay, az) = points[k].xyz, slice iz
 * fixes the probe height pz = az - iz*dh and relaxes (x, y) from the anchor by damped velocity
 * steps (v = v (1 - damp) + F dt; p += v dt) in the grid force plus the spring -k_lat (p - anchor),
 * at most ctl->max_iter steps, stopping when Fx^2 + Fy^2 < ctl->f2conv after a step
 * (ctl->check_before and ctl->damped are ignored).
 *   FEs (n_scan,nz,4): sample at the final position; disps (n_scan,nz,4) or NULL:
 *   (x - ax, y - ay, 0, 0); PPs or NULL: (x, y, pz, +/-iterations) as ppm_relax_strokes_tilted. */
int ppm_relax_strokes_2d(const float* grid, const PpmGridSpec* spec,
                         const float* points, int n_scan, int nz,
                         float k_lat, float dh, float dt, float damp, const PpmRelaxCtl* ctl,
                         float* FEs, float* disps, float* PPs, void* stream);

/* Unrelaxed sampling along strokes.
 * Replaces getFEinStrokesTilted (kernels/AFM.cl:210-233) as launched by
 * AFMulator.get_raw_FE (spammm/SPM/AFM.py:2523-2580).  tip supplies tipA/B/C;
 * dTip and dpos0 are host float[4]. */
int ppm_fe_in_strokes_tilted(const float* grid, const PpmGridSpec* spec,
                             const float* points, int n_scan, int nz,
                             const PpmTip* tip, const float dTip[4], const float dpos0[4],
                             float* FEs, void* stream);

/* Implicit-function adjoint of relaxed states (no counterpart on the reference's grid path;
 * it plays the role of pass 1 of the direct-Morse VJP, morseDirectStateAdjoint,
 * kernels/AFM.cl:1711-1845, for a trilinear-sampled grid).
 * At each stored probe position PPs[s].xyz the total force vanishes; for the upstream
 * cotangent dL_dFEs[s] (float4, w.r.t. the tip-rotated FE output of
 * ppm_relax_strokes_tilted) it solves J^T l = (dO/dq)^T u with the 3x3 Jacobian J of the
 * total force (FP64) and ADDS the resulting cotangent of every touched grid value into
 * grid_cot ((nz,ny,nx,4), caller-zeroed; float atomics).  lambdas (n_scan,nz,4): l and a
 * status word in .w.  A state is used only if it passes the gates of the reference's
 * direct-path adjoint (kernels/AFM.cl:1778, :1811-1840); status bits (0 = used): 1 residual
 * |G(q*)| >= 5e-4, 2 non-finite, 4 singular or condition number >= 1e6, 8 unstable (smallest
 * eigenvalue of the symmetric part of -J <= 1e-5).  A gated state adds nothing and has l = 0.
 * points/tip must be those of the forward call. */
int ppm_relax_adjoint(const float* grid, const PpmGridSpec* spec,
                      const float* points, int n_scan, int nz, const PpmTip* tip,
                      const float* PPs, const float* dL_dFEs,
                      float* grid_cot, float* lambdas, void* stream);

/* Per-state diagnostics of relaxed states for the trilinear grid model: the adjoint_diag record
 * of the reference's direct-path state adjoint (morseDirectStateAdjoint, kernels/AFM.cl:1774-1844)
 * on the grid path, without a cotangent.  diag (n_scan,nz,4) = (|G(q*)|, smallest eigenvalue of
 * the symmetric part of -J, infinity-norm condition number of J, status bits of
 * ppm_relax_adjoint: 1 residual, 2 non-finite, 4 singular/ill-conditioned, 8 unstable).  A small
 * smallest eigenvalue marks a probe near a fold of its equilibrium branch, where a small change
 * of the field moves it to another branch; callers exclude such probes from a loss before they
 * differentiate it.  points/tip/PPs as for ppm_relax_adjoint. */
int ppm_relax_state_diag(const float* grid, const PpmGridSpec* spec,
                         const float* points, int n_scan, int nz, const PpmTip* tip,
                         const float* PPs, float* diag, void* stream);

/* Newton polish of converged states (no reference counterpart; used by the differentiable
 * path so that the forward sits on the equilibrium the implicit-function adjoint assumes).
 * For every state with PPs[s].w > 0, up to `steps` FP64 Newton steps q <- q - J^{-1} G(q) on
 * the trilinear model (J as in ppm_relax_adjoint), each kept only if it lowers |G|; PPs.xyz
 * is updated in place and, if FEs is not NULL, FEs[s] is resampled there (tip-rotated).
 * points/tip must be those of the ppm_relax_strokes_tilted call that produced PPs. */
int ppm_relax_polish(const float* grid, const PpmGridSpec* spec,
                     const float* points, int n_scan, int nz, const PpmTip* tip,
                     int steps, float* PPs, float* FEs, void* stream);

/* Largest atom count of the gridless direct path (the reference's limit,
 * spammm/SPM/AFM.py:2105-2106: nAtoms in [1,128], no fallback). */
#define PPM_DIRECT_MAX_ATOMS 128

/* Gridless direct Morse + point-charge relaxed scan: relaxStrokesTiltedMorseDirect
 * (kernels/AFM.cl:1419-1511) as launched by AFMulator.run_scan_morse_direct
 * (spammm/SPM/AFM.py:2063-2233).  atoms (na,4) = (x,y,z,Q); cMs (na,4) = (R0,E0,K<0,0);
 * pair law uncapped, no periodic images; Qs/QZs = tip charges and their world-z offsets
 * (scaled by COULOMB_CONST inside).  tip->tipC.w = dtip (must be < 0).  The |f|^2 < f2conv
 * test runs before each update (ctl->check_before is ignored); ctl->max_iter plays the role of
 * N_RELAX_STEP_MAX.  FEs (n_scan,nz,4): tip-rotated sample FE at the final probe position;
 * PPs (n_scan,nz,4): (position, +iterations if converged else -iterations), may be NULL.
 * 1 <= na <= PPM_DIRECT_MAX_ATOMS. */
int ppm_relax_direct_morse(const float* atoms, const float* cMs, int na,
                           const float* points, int n_scan, int nz, const PpmTip* tip,
                           const float Qs[4], const float QZs[4], const PpmRelaxCtl* ctl,
                           float* FEs, float* PPs, void* stream);

/* Bytes of device workspace of ppm_direct_morse_vjp (per-block FP64 partial sums). */
size_t ppm_direct_morse_vjp_workspace_size(int na, int n_scan, int nz);

/* Implicit-equilibrium VJP of ppm_relax_direct_morse: morseDirectStateAdjoint,
 * morseDirectParamPartials and reduceMorseDirectParamPartials (kernels/AFM.cl:1711-2029) as
 * driven by AFMulator.vjp_scan_morse_direct (spammm/SPM/AFM.py:2241-2519), evaluated in FP64.
 * PPs / dL_dFEs (n_scan,nz,4) as produced / consumed around the forward call (same atoms,
 * points and tip).  Per state: lambdas (n_scan,nz,4) = (lambda, 0); diag (n_scan,nz,4) =
 * (residual, lambda_min, condition estimate, status) with the reference's status codes
 * 0 ok, 1 not converged or residual >= 5e-4, 2 non-finite, 3 unstable (lambda_min <= 1e-5),
 * 4 singular or condition >= 1e6; a state with status != 0 contributes nothing.
 * grad (na,6) double, theta order (x,y,z,R0,E0,Q), overwritten (no atomics: deterministic). */
int ppm_direct_morse_vjp(const float* atoms, const float* cMs, int na,
                         const float* points, int n_scan, int nz, const PpmTip* tip,
                         const float Qs[4], const float QZs[4],
                         const float* PPs, const float* dL_dFEs,
                         float* lambdas, float* diag, void* workspace, size_t workspace_bytes,
                         double* grad, void* stream);

/* Bytes of device workspace (work queue, per-molecule counters and flags) of a batch of nmol. */
size_t ppm_batch_workspace_size(int nmol);

/* Batched forward (no reference counterpart: the reference runs one image at a time through
 * make_forcefield / run_scan / compute_df, spammm/SPM/AFM.py:897-1010, :3048-3089).
 * Runs grid projection, relaxed scan (relaxStrokesTilted semantics) and the df operator of
 * every molecule described by `mols` (device array of params->nmol PpmMolDesc) in one
 * persistent launch of `workers` workgroups.
 *   atoms / cpar: concatenated rows of all molecules (PpmMolDesc.atom_off);
 *   grids: stored z windows of all grids; FEs, PPs (PPs may be NULL): (n_pix_total, nz, 4);
 *   df (may be NULL): (n_pix_total, df_nz_out), operator dfW (df_nz_out*df_taps), dfStart;
 *   workspace: >= ppm_batch_workspace_size(nmol) bytes; after the call workspace int[1] is
 *     the error word (bit 1: a wait hit its spin limit) and ints [4+nmol, 4+2*nmol) are
 *     per-molecule flags (nonzero: a probe left the stored z window — recompute that molecule
 *     with the whole grid);
 *   reset: 1 = clear queue, counters and flags first; 2 = clear the queue head only (second
 *     launch of a two-launch GRID_ONLY + RELAX_ONLY sequence); 0 = no clearing.
 * The caller builds the descriptors (item_off, n_grid_items, n_relax_items, window, offsets)
 * consistently with the buffers; the library cannot check device-resident descriptors. */
int ppm_batch_forward(const PpmMolDesc* mols, const PpmBatchParams* params,
                      const float* atoms, const float* cpar, float* grids,
                      float* FEs, float* PPs, float* df, const float* dfW, const int* dfStart,
                      void* workspace, size_t workspace_bytes, int workers, int reset,
                      void* stream);

/* Largest nmol of ppm_batch_forward_noalloc. */
#define PPM_BATCH_NOALLOC_MAX_MOL 65536

/* ppm_batch_forward without a workspace argument and without any allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * results as ppm_batch_forward, except that the work queue, per-molecule counters and window
 * flags live in a library-owned scratch array in device memory (one per GPU, 512 KiB, kept until
 * the process ends; its address is looked up on the first call on each device, which must be the
 * current device) instead of a caller workspace; params->nmol <= PPM_BATCH_NOALLOC_MAX_MOL.
 *   status: device int[1 + nmol], written after the launch: status[0] = error word (bit 1: a
 *     wait hit its spin limit), status[1 + m] = window flag of molecule m (nonzero: recompute m
 *     with the whole grid);
 *   reset: 1 (clear queue, counters and flags first) or 2 (clear the queue head only: the
 *     RELAX_ONLY launch that follows a GRID_ONLY launch); 0 is rejected.
 * Every call on one GPU shares that GPU's scratch array: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialized by the caller. */
int ppm_batch_forward_noalloc(const PpmMolDesc* mols, const PpmBatchParams* params,
                              const float* atoms, const float* cpar, float* grids,
                              float* FEs, float* PPs, float* df, const float* dfW, const int* dfStart,
                              int* status, int workers, int reset, void* stream);

/* Bytes of device workspace ppm_grid_adjoint needs for na atoms on this grid (0 on bad input). */
size_t ppm_grid_adjoint_workspace_size(int na, const PpmGridSpec* spec);

/* Grid -> atom adjoint of ppm_grid_project (no reference counterpart; the reference's only
 * adjoint is the gridless morseDirectParamPartials/reduce pair, kernels/AFM.cl:1850-2029).
 * Given grid_cot = dL/dgrid, writes grad (na,6) FP64 = dL/d(x, y, z, q, p0, p1) per atom,
 * with (p0,p1) = (c6,c12) for PPM_MODEL_LJ and (R0,E0) for PPM_MODEL_MORSE, summed over
 * periodic images.  grid is the forward output (used to chain through the |E|>100
 * rescale); the other arguments are those of the forward call.  workspace: device memory
 * of at least ppm_grid_adjoint_workspace_size(na, spec) bytes.  params: 1 = all six columns;
 * 0 = positions and charges only (columns 4, 5 written as 0; ~12 % less arithmetic). */
int ppm_grid_adjoint(int model, const float* atoms, const float* cpar, int na,
                     const PpmGridSpec* spec, const float Qs[4], const float QZs[4],
                     int npbc_a, int npbc_b, const float lvA[3], const float lvB[3], int cap,
                     const float* grid, const float* grid_cot, void* workspace,
                     size_t workspace_bytes, int params, double* grad, void* stream);

/* ppm_grid_adjoint for a grid made by ppm_grid_project_add: gradients of `atoms` only; `base`
 * (the same base grid) enters the chain rule of the |E|>100 rescale of capped voxels. */
int ppm_grid_adjoint_add(int model, const float* atoms, const float* cpar, int na,
                         const PpmGridSpec* spec, const float Qs[4], const float QZs[4],
                         int npbc_a, int npbc_b, const float lvA[3], const float lvB[3], int cap,
                         const float* base, const float* grid, const float* grid_cot,
                         void* workspace, size_t workspace_bytes, int params, double* grad, void* stream);

/* Linear operator along z, one output per (pixel, z_out):
 *   out[p*nz_out + j] = sum_{t<taps} W[j*taps + t] * in[(p*nz_in + start[j] + t)*in_stride + in_offset]
 * Replaces the host NumPy Fz -> df conversions compute_df / compute_df_amp /
 * get_df (spammm/SPM/AFM.py:3048-3089, :2591-2599); the host builds W/start for
 * the chosen definition (finite difference, finite-amplitude mean, Giessibl).
 *   W (nz_out*taps) float, start (nz_out) int: device arrays; indices start[j]+t are
 *   clamped into [0, nz_in-1]. */
int ppm_z_stencil(const float* in, int in_stride, int in_offset, int n_pix, int nz_in,
                  const float* W, const int* start, int taps, int nz_out,
                  float* out, void* stream);

/* General banded linear operator along one axis of a strided float volume:
 *   out[(o*n_axis_out + j)*n_inner + i] = alpha * sum_{t<taps} W[j*taps+t] * v(o, start[j]+t, i)
 *                                         (+ previous out value if accumulate != 0)
 *   v(o, m, i) = sum_{c<nch} chw[c] * in[o*in_outer_stride + m*in_axis_stride + i*in_inner_stride + c]
 * (strides in floats).  Used for the directional derivatives of compute_df_dir /
 * compute_df_amp_dir (spammm/SPM/AFM.py:3105-3123, :3203-3236), which project F onto
 * the oscillation direction (chw = n) and differentiate along x, y or z, and for the
 * transpose of the df operator in df_loss_seed (spammm/SPM/AFM.py:3125-3200).
 * Indices start[j]+t are clamped into [0, n_axis_in-1]. */
int ppm_axis_operator(const float* in, long long in_outer_stride, long long in_axis_stride,
                      long long in_inner_stride, const float chw[4], int nch,
                      int n_outer, int n_axis_in, int n_inner,
                      const float* W, const int* start, int taps, int n_axis_out,
                      float alpha, int accumulate, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PPM_API_H */
