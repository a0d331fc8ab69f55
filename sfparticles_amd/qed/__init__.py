"""QED model selection (sfparticles/qed/__init__.py).

As in the reference, ``SFPARTICLES_OPTICAL_DEPTH=1`` in the environment at
import time selects the optical-depth model; otherwise the per-step
rejection-sampling model is used.  ``set_model`` switches it before species
are created (species created under the optical-depth model carry ``tau``).

Random draws are Philox4x32-10 streams keyed by ``seed`` (environment
``SFPARTICLES_SEED``, default 0) with counter (particle id, step, species,
slot).
"""
from os import getenv

_use_optical_depth = getenv("SFPARTICLES_OPTICAL_DEPTH") == "1"
seed = int(getenv("SFPARTICLES_SEED", "0"))

# draws made by per-operation calls (Particles._photon_event etc.) use this
# slot bit so they never coincide with the draws of a simulation step
PER_OP_SLOT = 0x80


def set_model(model: str) -> None:
    """'rs' (rejection sampling) or 'od' (optical depth)."""
    global _use_optical_depth
    if model not in ("rs", "od"):
        raise ValueError("model must be 'rs' or 'od'")
    _use_optical_depth = model == "od"


def model() -> str:
    return "od" if _use_optical_depth else "rs"


def set_seed(value: int) -> None:
    global seed
    seed = int(value)


_tables_handle = None


def _forget_device_tables():
    """The next device_tables() uploads the tables anew (tables.use_tables)."""
    global _tables_handle
    _tables_handle = None


def device_tables():
    """Library handle of the RS and OD tables (uploaded once per process)."""
    global _tables_handle
    if _tables_handle is None:
        import ctypes as C
        from .. import _lib as L
        from .tables import load_rs_tables, load_od_tables
        rs = load_rs_tables()
        od = load_od_tables()
        rsc = L.RSTablesC(rs.int_Ai.ctypes.data, rs.Aip.ctypes.data, rs.z_N, rs.z_min, rs.z_max,
                          rs.z_delta, rs.Ai.ctypes.data)
        odc = L.ODTablesC(od.photon_total.ctypes.data, od.pair_total.ctypes.data,
                          od.photon_cdf.ctypes.data, od.pair_cdf.ctypes.data, od.chi_N,
                          od.log_chi_min, od.log_chi_max, od.log_chi_delta, od.delta_N,
                          od.log_delta_min, od.log_delta_max, od.log_delta_delta,
                          od.chi_grid.ctypes.data, od.delta_grid.ctypes.data, od.chi_min,
                          od.chi_max)
        h = C.c_void_p()
        L.call("sfp_tables_create", C.byref(rsc), C.byref(odc), C.byref(h))
        _tables_handle = h.value
    return _tables_handle
