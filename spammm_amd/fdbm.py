"""FDBM force-field producer on the device: FFT Pauli overlap, FFT Poisson / electrostatic
convolution, London dispersion, composition and the central-difference force grid that
`AFMulator.scan_fdbm` relaxes on.

Mirrors the reference's FDBM stage 3 (spammm/SPM/AFM.py:3304-3566, :3580-3587, :4085-4253,
:4362-4431; kernels/AFM.cl:1066-1113, :1151-1303): same function names, arguments, return
types, array layouts ((nx,ny,nz) C-order host arrays) and exceptions.  The FFTs are FP32
hipFFT real<->complex transforms (the reference uses complex64 clFFT); the product paths run on
the GPU and raise without it.  Device-level helpers take and return torch device tensors;
the reference-named functions take and return NumPy arrays.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import device as D

COULOMB_CONST = 14.3996448915                      # eV A / e^2 (kernels/common.cl:29)
POISSON_SCALE = 4.0 * np.pi * COULOMB_CONST
C6_DEFAULT = {1: 6.5, 6: 24.0, 7: 20.0, 8: 15.0}  # spammm/SPM/AFM.py:4380-4381


# ── host helpers (grid planning, tip placement) ──────────────────────────────────────

def is_fft_friendly(n):
    """Only prime factors 2, 3, 5, 7 (spammm/SPM/AFM.py:3355-3362)."""
    n = int(n)
    if n <= 0:
        return False
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def round_fft_friendly(n, multiple=8):
    """Round up to a multiple of `multiple` with only prime factors 2, 3, 5, 7 (:3364-3370)."""
    n = int(((int(n) + multiple - 1) // multiple) * multiple)
    while not is_fft_friendly(n):
        n += multiple
    return n


def plan_fft_friendly_grid(atomPos, step, margin, z_vac):
    """COM-centred xy, z-symmetric FFT-friendly grid around a molecule (spammm/SPM/AFM.py:3603-3646):
    returns (grid_spec dict, origin, ngrid, step); ValueError if a dimension is not friendly."""
    atomPos = np.asarray(atomPos, dtype=np.float64)
    step, margin, z_vac = float(step), float(margin), float(z_vac)
    com = atomPos.mean(axis=0)
    mol_z = float(atomPos[:, 2].mean())
    ext = [(float(atomPos[:, a].min()) - margin, float(atomPos[:, a].max()) + margin) for a in (0, 1)]
    ext.append((mol_z - z_vac, mol_z + z_vac))
    ngrid = tuple(round_fft_friendly(int(np.ceil((hi - lo) / step)) + 1) for lo, hi in ext)
    for dim, name in zip(ngrid, ("nx", "ny", "nz")):
        if not is_fft_friendly(dim):
            raise ValueError(f"plan_fft_friendly_grid: {name}={dim} is not clFFT-friendly "
                             f"(prime factors must be 2,3,5,7 only)")
    origin = np.array([com[0] - 0.5 * ngrid[0] * step, com[1] - 0.5 * ngrid[1] * step,
                       mol_z - 0.5 * ngrid[2] * step], dtype=np.float64)
    spec = {"origin": origin, "ngrid": ngrid, "dA": [step, 0.0, 0.0], "dB": [0.0, step, 0.0],
            "dC": [0.0, 0.0, step]}
    return spec, origin, ngrid, step


def pad_roll_shifts(co_rho, target_shape):
    """Offset that centres a raw tip density in the target grid and the roll that then moves its
    |peak| to index (0,0,0) (spammm/SPM/AFM.py:4178-4189): (ox, oy, oz, rx, ry, rz)."""
    co_rho = np.asarray(co_rho)
    off = [(int(t) - int(c)) // 2 for t, c in zip(target_shape, co_rho.shape)]
    peak = np.unravel_index(int(np.argmax(np.abs(co_rho))), co_rho.shape)
    return off[0], off[1], off[2], off[0] + int(peak[0]), off[1] + int(peak[1]), off[2] + int(peak[2])


# ── device-level operations (torch float32 device tensors, (nx,ny,nz) C-order) ─────────

def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _field(t, name):
    D._f32dev(t, name)
    if t.dim() != 3:
        raise ValueError(f"{name} must be a (nx,ny,nz) field, got {tuple(t.shape)}")
    return t


_release_at_exit = False


def release_plans():
    """Destroys the library's cached hipFFT plans (ppm_fdbm_release_plans); the next FFT recreates
    them.  Registered to run at interpreter exit after the first FFT."""
    _lib.check(_lib.load().ppm_fdbm_release_plans())


def fft_combine(mode, a, b, step, scale, out=None):
    """out = scale * IFFT(FFT(a) (op) FFT(b) [/ k^2]) for the PPM_FFT_* modes (normalised inverse)."""
    import torch
    _field(a, "a")
    if b is not None:
        _field(b, "b")
        if b.shape != a.shape:
            raise ValueError("a and b must have the same shape")
    nx, ny, nz = (int(v) for v in a.shape)
    out = torch.empty_like(a) if out is None else _field(out, "out")
    L = _lib.load()
    nb = int(L.ppm_fdbm_workspace_size(nx, ny, nz))
    if nb == 0:
        raise ValueError(f"invalid FFT grid {a.shape}")
    ws = torch.empty(nb, dtype=torch.uint8, device=a.device)
    _lib.check(L.ppm_fdbm_fft_combine(int(mode), nx, ny, nz, float(step), a.data_ptr(),
                                      b.data_ptr() if b is not None else None, float(scale), out.data_ptr(),
                                      ws.data_ptr(), nb, D._stream(a)))
    global _release_at_exit
    if not _release_at_exit:
        import atexit
        atexit.register(release_plans)
        _release_at_exit = True
    return out


def pauli_scale_(field, A, beta, clip=1e-30):
    """In place: field = A * max(field, clip)^beta."""
    _field(field, "field")
    _lib.check(_lib.load().ppm_fdbm_pauli_scale(field.data_ptr(), int(field.numel()), float(A), float(beta),
                                                float(clip), D._stream(field)))
    return field


def pad_roll(src, shape, offset, roll):
    """Pad `src` into a zero grid of `shape` at `offset`, then roll by -roll (dst[i] = padded[i+roll])."""
    import torch
    _field(src, "src")
    dst = torch.empty(tuple(int(v) for v in shape), dtype=torch.float32, device=src.device)
    _lib.check(_lib.load().ppm_fdbm_pad_roll(src.data_ptr(), ctypes.byref(_i3(src.shape)), dst.data_ptr(),
                                             ctypes.byref(_i3(shape)), ctypes.byref(_i3(offset)),
                                             ctypes.byref(_i3(roll)), D._stream(src)))
    return dst


def flip3(src):
    """dst[ix,iy,iz] = src[nx-1-ix, ny-1-iy, nz-1-iz]."""
    import torch
    _field(src, "src")
    dst = torch.empty_like(src)
    _lib.check(_lib.load().ppm_fdbm_flip3(src.data_ptr(), dst.data_ptr(), ctypes.byref(_i3(src.shape)),
                                          D._stream(src)))
    return dst


def dispersion(atoms_w, origin, step, shape, RA=1.5, out=None, accumulate=False):
    """out (+)= -sum_a w_a / (|p - r_a|^2 + RA^2)^3 on p = origin + step*(ix,iy,iz);
    atoms_w (na,4) float32 device = (x, y, z, sqrt(C6_a C6_tip))."""
    import torch
    D._f32dev(atoms_w, "atoms_w", 4)
    shape = tuple(int(v) for v in shape)
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs an output field")
        out = torch.empty(shape, dtype=torch.float32, device=atoms_w.device)
    else:
        _field(out, "out")
    org = (ctypes.c_float * 3)(*[float(v) for v in np.asarray(origin, float).ravel()[:3]])
    _lib.check(_lib.load().ppm_fdbm_dispersion(atoms_w.data_ptr(), int(atoms_w.shape[0]), ctypes.byref(org),
                                               float(step), ctypes.byref(_i3(shape)), float(RA), int(bool(accumulate)),
                                               out.data_ptr(), D._stream(atoms_w)))
    return out


def compose_gradient(fields, step, grid=None):
    """Force grid (nz,ny,nx,4) = (-grad E, E) of E = sum(fields), periodic central differences."""
    import torch
    fields = [f for f in fields if f is not None]
    if not 1 <= len(fields) <= 3:
        raise ValueError("one to three energy fields")
    for k, f in enumerate(fields):
        _field(f, f"field{k}")
        if f.shape != fields[0].shape:
            raise ValueError("energy fields must have the same shape")
    nx, ny, nz = (int(v) for v in fields[0].shape)
    if grid is None:
        grid = torch.empty((nz, ny, nx, 4), dtype=torch.float32, device=fields[0].device)
    else:
        D._f32dev(grid, "grid", 4)
        if grid.numel() != 4 * nx * ny * nz:
            raise ValueError("grid has the wrong size")
    ptr = [f.data_ptr() for f in fields] + [None] * (3 - len(fields))
    _lib.check(_lib.load().ppm_fdbm_compose_gradient(ptr[0], ptr[1], ptr[2], ctypes.byref(_i3((nx, ny, nz))),
                                                     float(step), grid.data_ptr(), D._stream(fields[0])))
    return grid


# ── reference-named functions (NumPy in, NumPy out) ─────────────────────────────────

def _up(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=D.require_gpu())


def _iso_step(step):
    if np.ndim(step) != 0:
        s = np.asarray(step, dtype=np.float64).ravel()[:3]
        if not np.allclose(s, s[0], rtol=0, atol=0):
            raise ValueError("the device FFT path needs an isotropic step")
        return float(s[0])
    return float(step)


def _grads(E, step):
    return np.stack([np.gradient(E, step, axis=i) for i in range(3)], axis=-1)


def fft_poisson(rho, step):
    """V(r) [eV/e] from a charge density rho(r) [e/A^3]: V_k = 4 pi C rho_k / k^2, V_0 = 0
    (spammm/SPM/AFM.py:3580-3587, :3280-3298)."""
    st = _iso_step(step)
    return fft_combine(_lib.PPM_FFT_POISSON, _up(rho), None, st, POISSON_SCALE).cpu().numpy()


def compute_pauli_overlap(rho_grid, rho_tip_total, step, tip_rolled=False):
    """overlap(R) = dV * IFFT(FFT(rho) . conj(FFT(tip))), clipped at 1e-30 (spammm/SPM/AFM.py:4092-4103).
    `tip_rolled` is accepted for the reference's signature; the tip is used as given."""
    st = _iso_step(step)
    out = fft_combine(_lib.PPM_FFT_CORRELATE, _up(rho_grid), _up(rho_tip_total), st, st ** 3)
    return out.clamp_(min=1e-30).cpu().numpy()


def scale_pauli_field(overlap_raw, step, A_pauli, beta_pauli, return_grads=True):
    """E_pauli = A * overlap^beta (spammm/SPM/AFM.py:4106-4128); grads as np.gradient."""
    E = pauli_scale_(_up(overlap_raw), A_pauli, beta_pauli, 1e-30).cpu().numpy()
    return (E, _grads(E, step)) if return_grads else E


def _tip_kernel(rho_tip_delta, tip_rolled):
    """The reversed tip of the electrostatic convolution (spammm/SPM/AFM.py:4137-4175)."""
    t = _up(rho_tip_delta)
    k = flip3(t)
    if not tip_rolled:
        nx, ny, nz = (int(v) for v in k.shape)
        k = pad_roll(k, k.shape, (0, 0, 0), (nx // 2, ny // 2, nz // 2))
    return k


def compute_es_conv_field(V_ES, rho_tip_delta, step, tip_rolled=False, return_grads=True):
    """E(R) = dV * IFFT(FFT(V) . FFT(reversed tip)) (spammm/SPM/AFM.py:4153-4175)."""
    st = _iso_step(step)
    E = fft_combine(_lib.PPM_FFT_CONVOLVE, _up(V_ES), _tip_kernel(rho_tip_delta, tip_rolled), st,
                    st ** 3).cpu().numpy()
    return (E, _grads(E, st)) if return_grads else E


def compute_es_fused_field(rho_diff, rho_tip_delta, step, tip_rolled=True, return_grads=False):
    """Poisson and the electrostatic convolution in one spectrum product (spammm/SPM/AFM.py:4192-4201)."""
    st = _iso_step(step)
    E = fft_combine(_lib.PPM_FFT_POISSON_CONVOLVE, _up(rho_diff), _tip_kernel(rho_tip_delta, tip_rolled), st,
                    POISSON_SCALE * st ** 3).cpu().numpy()
    return (E, _grads(E, st)) if return_grads else E


def _c6_atoms(atomPos, atomTypes, C6_atom_dict, C6_CO, default):
    table = C6_DEFAULT if C6_atom_dict is None else C6_atom_dict
    atomPos = np.asarray(atomPos, dtype=np.float64)
    a = np.zeros((len(atomPos), 4), dtype=np.float32)
    a[:, :3] = atomPos[:, :3]
    a[:, 3] = np.sqrt(np.array([table.get(z, default) for z in np.asarray(atomTypes).tolist()], np.float32)
                      * np.float32(C6_CO))
    return a


def compute_dispersion_grid(atomPos, atomTypes, origin, step, ngrid, C6_atom_dict=None, C6_CO=30.0, RA=1.5,
                            use_opencl=True, return_grads=True, afmulator=None):
    """E_vdw = -sum_a sqrt(C6_a C6_CO) / (r^2 + RA^2)^3 on origin + step*(ix,iy,iz)
    (spammm/SPM/AFM.py:4362-4431; unknown atom types take C6 = 1).  `use_opencl` and
    `afmulator` are accepted for the reference's signature; the grid is computed on the GPU."""
    a = _c6_atoms(atomPos, atomTypes, C6_atom_dict, C6_CO, 1.0)
    E = dispersion(_up(a), origin, float(step), tuple(int(v) for v in ngrid[:3]), RA).cpu().numpy()
    return (E, _grads(E, float(step))) if return_grads else E


def compute_vdw_field(atomPos, atomTypes, origin, step, ngrid, C6_table=None, C6_CO=30.0, RA=1.5):
    """The same dispersion field with unknown types at C6 = 10 and gradients
    (spammm/SPM/AFM.py:4256-4289): returns (E_vdw, grads)."""
    a = _c6_atoms(atomPos, atomTypes, C6_table, C6_CO, 10.0)
    E = dispersion(_up(a), origin, float(step), tuple(int(v) for v in ngrid[:3]), RA).cpu().numpy()
    return E, _grads(E, float(step))


def _place_tip(raw, target_shape, already_rolled):
    raw = np.asarray(raw, dtype=np.float32)
    if already_rolled or tuple(raw.shape) == tuple(target_shape):
        if tuple(raw.shape) != tuple(target_shape):
            raise ValueError(f"rolled tip has shape {raw.shape}, grid {tuple(target_shape)}")
        return _up(raw)
    ox, oy, oz, rx, ry, rz = pad_roll_shifts(raw, target_shape)
    return pad_roll(_up(raw), target_shape, (ox, oy, oz), (rx, ry, rz))


def stage3_fdbm_fields_fast(afmulator, rho_scf, rho_diff, tip_total_raw, tip_delta_raw, origin, step, ngrid,
                            atomPos, atomTypes, A_pauli, beta_pauli, C6_CO=30.0, tip_already_rolled=False,
                            download_fields=True):
    """FDBM stage 3 on the device (spammm/SPM/AFM.py:4204-4253): Pauli overlap scaled A.o^beta,
    fused Poisson + electrostatic convolution, London dispersion, composition and the force grid,
    installed as `afmulator`'s FDBM grid without a host round trip.  Returns
    (V_ES=None, E_pauli, E_ES, E_vdw, F_total (nx,ny,nz,4)); the three energies only if
    `download_fields`."""
    target = tuple(int(x) for x in ngrid[:3])
    st = _iso_step(step)
    dV = st ** 3
    rho = _up(rho_scf)
    rdiff = _up(rho_diff)
    if tuple(rho.shape) != target or tuple(rdiff.shape) != target:
        raise ValueError(f"densities must have the grid shape {target}")
    tip_tot = _place_tip(tip_total_raw, target, tip_already_rolled)
    tip_del = _place_tip(tip_delta_raw, target, tip_already_rolled)
    E_p = pauli_scale_(fft_combine(_lib.PPM_FFT_CORRELATE, rho, tip_tot, st, dV), A_pauli, beta_pauli, 1e-30)
    E_e = fft_combine(_lib.PPM_FFT_POISSON_CONVOLVE, rdiff, flip3(tip_del), st, POISSON_SCALE * dV)
    E_v = dispersion(_up(_c6_atoms(atomPos, atomTypes, None, C6_CO, 1.0)), origin, st, target)
    grid = compose_gradient((E_p, E_e, E_v), st)
    afmulator.install_fdbm_grid(grid, origin, st, target)
    F_total = np.ascontiguousarray(grid.cpu().numpy().transpose(2, 1, 0, 3))
    if download_fields:
        return None, E_p.cpu().numpy(), E_e.cpu().numpy(), E_v.cpu().numpy(), F_total
    return None, None, None, None, F_total
