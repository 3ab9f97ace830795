"""Element parameters and the tip-sample mixing rules of the probe-particle model.

Mixing (AFMulator.assign_params, spammm/SPM/AFM.py:730-768; _morse_atoms_from_Z,
spammm/SPM/AFM_utils.py:3993-4017):
    R0_ij = tip_R + RvdW_i,   E0_ij = sqrt(|tip_E * EvdW_i|)
    LJ:    (c6, c12) = (2 E0 R0^6, E0 R0^12)          (getLJ, kernels/AFM.cl:738-745)
    Morse: (R0, E0, K=tip_alpha<0, 0)                  (getMorse, kernels/AFM.cl:747-757)

``read_element_types`` parses the ElementTypes.dat column layout used by the
reference (spammm/topology/FFparams.py:103-143: name iZ ne nv npi color Rcov RvdW
EvdW ...); only the columns the mixing rules need are kept.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

N_M_PER_EV_A2 = 16.02176634          # spammm/SPM/AFM.py:67-72
EV_A2_PER_N_M = 1.0 / N_M_PER_EV_A2


def stiffness_Nm_to_eVA2(k_Nm: float) -> float:
    return float(k_Nm) * EV_A2_PER_N_M


def stiffness_eVA2_to_Nm(k_eVA2: float) -> float:
    return float(k_eVA2) * N_M_PER_EV_A2


@dataclass(frozen=True)
class ElementType:
    name: str
    iZ: int
    RvdW: float
    EvdW: float


# Values of data/ElementTypes.dat (H :4, C :9, N :10, O :11) for the elements the
# synthetic workloads use, and the C/H values the benchmark configuration states
# explicitly (C R0=1.908 A E0=0.0037 eV, H R0=1.487 A E0=0.00068 eV).
ELEMENTS_REFERENCE = {
    "H": ElementType("H", 1, 1.4430, 0.00190802059),
    "C": ElementType("C", 6, 1.9255, 0.00455323095),
    "N": ElementType("N", 7, 1.8300, 0.00299212319),
    "O": ElementType("O", 8, 1.7500, 0.00260184625),
    "Na": ElementType("Na", 11, 1.4915, 0.00130092313),     # data/ElementTypes.dat:14
    "Cl": ElementType("Cl", 17, 1.9735, 0.00984365166),     # data/ElementTypes.dat:21
}
ELEMENTS_BENCH = {
    "H": ElementType("H", 1, 1.487, 0.00068),
    "C": ElementType("C", 6, 1.908, 0.0037),
    "N": ElementType("N", 7, 1.8300, 0.00299212319),
    "O": ElementType("O", 8, 1.7500, 0.00260184625),
    "Na": ElementType("Na", 11, 1.4915, 0.00130092313),
    "Cl": ElementType("Cl", 17, 1.9735, 0.00984365166),
}
Z_TO_NAME = {1: "H", 6: "C", 7: "N", 8: "O", 11: "Na", 17: "Cl"}


def read_element_types(path: str) -> dict:
    out = {}
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            parts = line.split()
            try:
                out[parts[0]] = ElementType(parts[0], int(parts[1]), float(parts[7]), float(parts[8]))
            except (IndexError, ValueError):
                continue
    return out


def mix_lj(enames, elements, tip_R: float, tip_E: float) -> np.ndarray:
    """(na,2) float32 LJ (c6, c12) rows."""
    c = np.zeros((len(enames), 2), dtype=np.float32)
    for i, e in enumerate(enames):
        et = elements.get(e)
        if et is None:
            raise KeyError(f"Element '{e}' has no parameters")
        R0 = tip_R + et.RvdW
        E0 = np.sqrt(abs(tip_E * et.EvdW))
        c[i] = [2.0 * E0 * R0 ** 6, E0 * R0 ** 12]
    return c


def mix_morse(enames, elements, tip_R: float, tip_E: float, tip_alpha: float) -> np.ndarray:
    """(na,4) float32 Morse (R0, E0, K, 0) rows."""
    c = np.zeros((len(enames), 4), dtype=np.float32)
    for i, e in enumerate(enames):
        et = elements.get(e)
        if et is None:
            raise KeyError(f"Element '{e}' has no parameters")
        c[i] = [tip_R + et.RvdW, np.sqrt(abs(tip_E * et.EvdW)), tip_alpha, 0.0]
    return c


def mix_R0E0(enames, elements, tip_R: float, tip_E: float):
    """Per-atom mixed (R0, E0) in float64 (the LJ parameters differentiated by the adjoint)."""
    R0 = np.array([tip_R + elements[e].RvdW for e in enames], dtype=np.float64)
    E0 = np.array([np.sqrt(abs(tip_E * elements[e].EvdW)) for e in enames], dtype=np.float64)
    return R0, E0


def morse_atoms_from_Z(atomPos, atomTypes, params_path=None, tip_R=1.452, tip_E=0.0006808, tip_alpha=-1.8):
    """Atom rows and Morse pair rows from positions and atomic numbers (_morse_atoms_from_Z,
    spammm/SPM/AFM_utils.py:3993-4017): atoms (na,4) float32 = (x, y, z, 0) and cMs (na,4) float32
    = the Morse mixing rule of the element with each atomic number.  The element table is
    params_path (ElementTypes.dat layout) or, without it, the built-in values of the reference's
    data/ElementTypes.dat.  An atomic number without an entry raises KeyError."""
    table = read_element_types(params_path) if params_path is not None else ELEMENTS_REFERENCE
    by_z = {int(et.iZ): name for name, et in table.items()}
    pos = np.asarray(atomPos, dtype=np.float32)
    zs = np.asarray(atomTypes).reshape(-1)[:len(pos)].astype(np.int64)
    missing = sorted({int(z) for z in zs if int(z) not in by_z})
    if missing:
        raise KeyError(f"morse_atoms_from_Z: no element entry for Z={missing}")
    atoms = np.zeros((len(pos), 4), dtype=np.float32)
    atoms[:, :3] = pos[:, :3]
    return atoms, mix_morse([by_z[int(z)] for z in zs], table, tip_R, tip_E, tip_alpha)

