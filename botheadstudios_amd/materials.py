"""Tillotson parameter sets and the 48-byte `SphEos` record.

Values are the reference catalogue's `tillotson` blocks (integrity-engine `data/materials.json`, read by
`crates/engine/src/materials.rs:437` and `eos.rs:117-171`). Field order is `eos.rs:27-47` /
`SphEos` (`gpu_sph.rs:42-55`): rho0, a, b, A, B, E0, E_iv, E_cv, alpha, beta (SI units).
"""
import numpy as np

# data/materials.json "granite" / "iron" / "basalt" tillotson blocks.
GRANITE = (2680.0, 0.5, 1.3, 1.8e10, 1.8e10, 1.6e7, 3.5e6, 1.8e7, 5.0, 5.0)
IRON = (7850.0, 0.5, 1.28, 1.28e11, 1.815e11, 1.425e7, 2.4e6, 8.67e6, 5.0, 5.0)
BASALT = (2700.0, 0.5, 1.5, 2.67e10, 2.67e10, 4.87e8, 4.72e6, 1.82e7, 5.0, 5.0)
PERIDOTITE = (3500.0, 0.5, 1.4, 1.31e11, 4.9e10, 5.5e8, 4.5e6, 1.45e7, 5.0, 5.0)
WATER = (998.0, 0.7, 0.15, 2.18e9, 1.325e10, 7.0e6, 4.19e5, 2.69e6, 10.0, 5.0)
ICE = (917.0, 0.3, 0.1, 9.47e9, 9.47e9, 1.0e7, 7.73e5, 3.04e6, 10.0, 5.0)

CATALOGUE = {
    "granite": GRANITE,
    "basalt": BASALT,
    "peridotite": PERIDOTITE,
    "iron": IRON,
    "water": WATER,
    "ice": ICE,
}

# data/materials.json "thermal.specific_heat" (J/kg/K): u = c T (hydrostatic.rs:236-242).
SPECIFIC_HEAT = {
    "granite": 790.0,
    "basalt": 840.0,
    "peridotite": 1000.0,
    "iron": 449.0,
    "water": 4186.0,
    "ice": 2090.0,
}

# Material indices used by the impact scenes (index into the EOS table passed to upload).
MAT_GRANITE = 0
MAT_IRON = 1

EOS_DTYPE = np.dtype([(k, "<f4") for k in
                      ("rho0", "a", "b", "cap_a", "cap_b", "e0", "e_iv", "e_cv", "alpha", "beta", "_p0", "_p1")])
assert EOS_DTYPE.itemsize == 48


def eos_table(names=("granite", "iron")):
    """SphEos records (f32) for the named materials, in order (gpu_sph.rs:99-117 from_tillotson)."""
    out = np.zeros(len(names), dtype=EOS_DTYPE)
    for i, n in enumerate(names):
        vals = CATALOGUE[n]
        for k, v in zip(EOS_DTYPE.names[:10], vals):
            out[i][k] = v
    return out


def eos_f64(names=("granite", "iron")):
    """The same table as f64 rows (10 parameters each), for the f64 CPU path."""
    return np.array([CATALOGUE[n] for n in names], dtype=np.float64)
