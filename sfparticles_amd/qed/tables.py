"""Generators and loaders for the QED rate tables.

Both generators restate the reference's published recipe:

* rejection-sampling tables (sfparticles/qed/rejection_sampling_tables.py:91-112):
  ``int_Ai(z) = quad(Ai, z, inf)`` and ``Ai'(z)`` from ``scipy.special.airy`` on
  ``z = linspace(0, 100, 1024)``;
* optical-depth tables (sfparticles/qed/optical_depth_tables.py:126-230):
  total rates ``quad(P, 0, 1)`` on ``chi = logspace(-3, 2, 256)``, and the
  cumulative rate along ``delta = logspace(-6, 0, 1024)`` (a quad up to the
  first grid point, then a right Riemann sum), for photon emission
  ``P = -alpha m c^2/hbar [int_Ai(z) + (2/z + chi_g sqrt z) Ai'(z)]``,
  ``z = (chi_g / chi_e / (chi_e - chi_g))^(2/3)`` and pair creation
  ``P = alpha m c^2/hbar [int_Ai(z) + (2/z - chi_g sqrt z) Ai'(z)]``.

The generated arrays ship with the package as ``.npy`` files (``tables/``);
``python sfparticles_amd/qed/tables.py`` regenerates them (about 30 s on 8
cores); the shipped files were generated with scipy 1.7.1.  The reference's own rejection-sampling HDF5 table is reproduced bit for
bit by ``rs_table_gen`` (checked by the test-suite against a fixture extracted
from it); the reference's optical-depth HDF5 blob is not available, so those
tables are pinned by the reference's known-answer vector for the photon total
rate and by closed-form checks.
"""
import multiprocessing
import os
import warnings

import numpy as np

if __package__:
    from .. import constants as K
else:  # run as a script, e.g. under an interpreter that lacks the package deps
    import importlib.util as _ilu

    _spec = _ilu.spec_from_file_location(
        "_sfp_constants", os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "constants.py"))
    K = _ilu.module_from_spec(_spec)
    _spec.loader.exec_module(K)

TABLE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables")

# grid definitions (sfparticles/qed/rejection_sampling_tables.py:91-94,
# sfparticles/qed/optical_depth_tables.py:201-204)
RS_Z_N, RS_Z_MIN, RS_Z_MAX = 1024, 0.0, 100.0
OD_CHI_N, OD_LOG_CHI_MIN, OD_LOG_CHI_MAX = 256, -3.0, 2.0
OD_DELTA_N, OD_LOG_DELTA_MIN = 1024, -6.0


# Ai and Ai' underflow to 0 for z > ~105; scipy returns NaN from z ~ 1e8 on,
# which the reference's recipe would carry into whole rows of the pair
# cumulative table (chi < 0.4).  Beyond Z_ZERO both are returned as 0.
Z_ZERO = 150.0


def _Ai(z):
    from scipy.special import airy
    return np.where(np.asarray(z) > Z_ZERO, 0.0, airy(np.minimum(z, Z_ZERO))[0])[()]


def _Aip(z):
    from scipy.special import airy
    return np.where(np.asarray(z) > Z_ZERO, 0.0, airy(np.minimum(z, Z_ZERO))[1])[()]


def int_Ai(z):
    """Integral of Ai from z to infinity (0 beyond Z_ZERO, see above)."""
    from scipy.integrate import quad
    if z > Z_ZERO:
        return 0.0
    return quad(_Ai, z, np.inf)[0]


def rs_table_gen(z_N=RS_Z_N, z_min=RS_Z_MIN, z_max=RS_Z_MAX, processes=None):
    """Return (int_Ai[z_N], Aip[z_N], attrs) on z = linspace(z_min, z_max, z_N)."""
    z = np.linspace(z_min, z_max, z_N)
    with multiprocessing.Pool(processes) as pool:
        ia = np.asarray(pool.map(int_Ai, z), dtype=np.float64)
    ap = np.asarray(_Aip(z), dtype=np.float64)
    attrs = dict(z_N=z_N, z_range=(z_min, z_max), z_delta=(z_max - z_min) / (z_N - 1))
    return ia, ap, attrs


def rs_ai_gen(z_N=RS_Z_N, z_min=RS_Z_MIN, z_max=RS_Z_MAX):
    """Ai(z) on the rejection-sampling grid: the spin term of the emission
    rate, -alpha m c^2/hbar * sigma * delta * Ai(z) / sqrt(z) (no counterpart
    in the reference)."""
    z = np.linspace(z_min, z_max, z_N)
    return np.asarray(_Ai(z), dtype=np.float64)


def photon_rate_for_delta(chi_e):
    """dW/d(delta) per unit proper time for photon emission at fixed chi_e."""
    factor = -K.alpha * K.m_e * K.c ** 2 / K.hbar

    def prob_(delta):
        if delta == 1.0:
            return 0.0
        chi_gamma = delta * chi_e
        chi_ep = chi_e - chi_gamma
        z = (chi_gamma / chi_e / chi_ep) ** (2 / 3)
        return factor * (int_Ai(z) + (2.0 / z + chi_gamma * np.sqrt(z)) * _Aip(z))

    return prob_


def pair_rate_for_delta(chi_gamma):
    """dW/d(delta) for Breit-Wheeler pair creation at fixed chi_gamma."""
    factor = K.alpha * K.m_e * K.c ** 2 / K.hbar

    def prob_(delta):
        if delta == 1.0:
            return 0.0
        chi_e = delta * chi_gamma
        chi_ep = chi_gamma - chi_e
        z = (chi_gamma / chi_e / chi_ep) ** (2 / 3)
        return factor * (int_Ai(z) + (2.0 / z - chi_gamma * np.sqrt(z)) * _Aip(z))

    return prob_


def photon_total_rate(chi_e):
    from scipy.integrate import quad
    return quad(photon_rate_for_delta(chi_e), 0, 1)[0]


def pair_total_rate(chi_gamma):
    from scipy.integrate import quad
    return quad(pair_rate_for_delta(chi_gamma), 0, 1)[0]


def _cumulative(P, delta_N, log_delta_min):
    from scipy.integrate import quad
    delta = np.logspace(log_delta_min, 0, delta_N)
    integ = np.zeros(delta_N)
    integ[0] = quad(P, 0, delta[0])[0]
    for i in range(1, delta_N):
        integ[i] = integ[i - 1] + P(delta[i]) * (delta[i] - delta[i - 1])
    return integ


def _photon_cumulative(args):
    chi, delta_N, log_delta_min = args
    return _cumulative(photon_rate_for_delta(chi), delta_N, log_delta_min)


def _pair_cumulative(args):
    chi, delta_N, log_delta_min = args
    return _cumulative(pair_rate_for_delta(chi), delta_N, log_delta_min)


def photon_total_table(chi_N=OD_CHI_N, log_chi_min=OD_LOG_CHI_MIN, log_chi_max=OD_LOG_CHI_MAX,
                       processes=None):
    chi = np.logspace(log_chi_min, log_chi_max, chi_N)
    with multiprocessing.Pool(processes) as pool:
        return np.asarray(pool.map(photon_total_rate, chi))


def od_table_gen(chi_N=OD_CHI_N, log_chi_min=OD_LOG_CHI_MIN, log_chi_max=OD_LOG_CHI_MAX,
                 delta_N=OD_DELTA_N, log_delta_min=OD_LOG_DELTA_MIN, processes=None):
    chi = np.logspace(log_chi_min, log_chi_max, chi_N)
    args = [(float(x), delta_N, log_delta_min) for x in chi]
    with multiprocessing.Pool(processes) as pool:
        photon_total = np.asarray(pool.map(photon_total_rate, chi))
        pair_total = np.asarray(pool.map(pair_total_rate, chi))
        photon_cdf = np.asarray(pool.map(_photon_cumulative, args))
        pair_cdf = np.asarray(pool.map(_pair_cumulative, args))
    attrs = dict(
        chi_N=chi_N, log_chi_range=(log_chi_min, log_chi_max),
        log_chi_delta=(log_chi_max - log_chi_min) / (chi_N - 1),
        delta_N=delta_N, log_delta_range=(log_delta_min, 0.0),
        log_delta_delta=(0 - log_delta_min) / (delta_N - 1),
    )
    return photon_total, pair_total, photon_cdf, pair_cdf, attrs


# ----------------------------------------------------------------- loading --
class RSTables:
    """Rejection-sampling tables as the kernels consume them."""

    def __init__(self, int_Ai, Aip, z_N, z_min, z_max, z_delta, Ai=None):
        self.int_Ai = np.ascontiguousarray(int_Ai, dtype=np.float64)
        self.Aip = np.ascontiguousarray(Aip, dtype=np.float64)
        if Ai is None:
            Ai = rs_ai_gen(int(z_N), float(z_min), float(z_max))
        self.Ai = np.ascontiguousarray(Ai, dtype=np.float64)
        self.z_N = int(z_N)
        self.z_min = float(z_min)
        self.z_max = float(z_max)
        self.z_delta = float(z_delta)


class ODTables:
    """Optical-depth tables plus the grids derived from their attributes."""

    def __init__(self, photon_total, pair_total, photon_cdf, pair_cdf, chi_N, log_chi_min,
                 log_chi_max, log_chi_delta, delta_N, log_delta_min, log_delta_max,
                 log_delta_delta):
        self.photon_total = np.ascontiguousarray(photon_total, dtype=np.float64)
        self.pair_total = np.ascontiguousarray(pair_total, dtype=np.float64)
        self.photon_cdf = np.ascontiguousarray(photon_cdf, dtype=np.float64)
        self.pair_cdf = np.ascontiguousarray(pair_cdf, dtype=np.float64)
        self.chi_N = int(chi_N)
        self.log_chi_min = float(log_chi_min)
        self.log_chi_max = float(log_chi_max)
        self.log_chi_delta = float(log_chi_delta)
        self.delta_N = int(delta_N)
        self.log_delta_min = float(log_delta_min)
        self.log_delta_max = float(log_delta_max)
        self.log_delta_delta = float(log_delta_delta)
        # grid points as the reference forms them: 10**(lo + i*step)
        # (sfparticles/qed/optical_depth_tables.py:76-77,107-108)
        self.chi_grid = np.array([10 ** (self.log_chi_min + i * self.log_chi_delta)
                                  for i in range(self.chi_N)], dtype=np.float64)
        self.delta_grid = np.array([10 ** (self.log_delta_min + j * self.log_delta_delta)
                                    for j in range(self.delta_N)], dtype=np.float64)
        # sfparticles/qed/optical_depth.py:14
        self.chi_min = 10.0 ** self.log_chi_min
        self.chi_max = 10.0 ** self.log_chi_max


_rs_cache = None
_od_cache = None


def load_rs_tables():
    global _rs_cache
    if _rs_cache is None:
        d = np.load(os.path.join(TABLE_DIR, "rs_tables.npz"))
        _rs_cache = RSTables(d["int_Ai"], d["Aip"], d["z_N"], d["z_range"][0], d["z_range"][1],
                             d["z_delta"], d["Ai"] if "Ai" in d else None)
    return _rs_cache


def load_od_tables():
    global _od_cache
    if _od_cache is None:
        d = np.load(os.path.join(TABLE_DIR, "od_tables.npz"))
        _od_cache = ODTables(d["photon_prob_rate_total"], d["pair_prob_rate_total"],
                             d["integral_photon_prob_along_delta"],
                             d["integral_pair_prob_along_delta"], d["chi_N"],
                             d["log_chi_range"][0], d["log_chi_range"][1], d["log_chi_delta"],
                             d["delta_N"], d["log_delta_range"][0], d["log_delta_range"][1],
                             d["log_delta_delta"])
    return _od_cache


def _h5py():
    try:
        import h5py
    except ImportError as exc:  # not a dependency of this package
        raise ImportError("reading the reference's HDF5 table files needs h5py") from exc
    return h5py


def read_rs_tables_h5(path):
    """The reference's rejection-sampling table file (written by
    qed/rejection_sampling_tables.py:96-112, read there at :16-24): datasets
    int_Ai and Aip, each carrying the attributes z_N, z_range, z_delta.  Ai
    (the spin term) is not in that file and is generated on the same grid."""
    h5py = _h5py()
    with h5py.File(path, "r") as f:
        d = f["int_Ai"]
        z_range = d.attrs["z_range"]
        return RSTables(d[()], f["Aip"][()], d.attrs["z_N"], z_range[0], z_range[1],
                        d.attrs["z_delta"])


def read_od_tables_h5(path):
    """The reference's optical-depth table file (written by
    qed/optical_depth_tables.py:206-230, read there at :17-30): datasets
    photon_prob_rate_total, pair_prob_rate_total,
    integral_photon_prob_along_delta, integral_pair_prob_along_delta; the
    attributes chi_N, log_chi_range, log_chi_delta, delta_N, log_delta_range,
    log_delta_delta on the file."""
    h5py = _h5py()
    with h5py.File(path, "r") as f:
        a = f.attrs
        return ODTables(f["photon_prob_rate_total"][()], f["pair_prob_rate_total"][()],
                        f["integral_photon_prob_along_delta"][()],
                        f["integral_pair_prob_along_delta"][()], a["chi_N"],
                        a["log_chi_range"][0], a["log_chi_range"][1], a["log_chi_delta"],
                        a["delta_N"], a["log_delta_range"][0], a["log_delta_range"][1],
                        a["log_delta_delta"])


def use_tables(rs=None, od=None):
    """Tables for the simulations created from now on (RSTables / ODTables,
    e.g. from read_rs_tables_h5); the default is the shipped pair.  Tables
    already on the device stay there for the simulations that use them."""
    global _rs_cache, _od_cache
    if rs is not None:
        _rs_cache = rs
    if od is not None:
        _od_cache = od
    from . import _forget_device_tables
    _forget_device_tables()


def main():
    warnings.filterwarnings("ignore")  # quad's IntegrationWarnings near z -> inf, as in the reference
    os.makedirs(TABLE_DIR, exist_ok=True)
    ia, ap, a = rs_table_gen()
    np.savez(os.path.join(TABLE_DIR, "rs_tables.npz"), int_Ai=ia, Aip=ap, z_N=a["z_N"],
             z_range=np.asarray(a["z_range"]), z_delta=a["z_delta"],
             Ai=rs_ai_gen(RS_Z_N, RS_Z_MIN, RS_Z_MAX))
    print("rejection-sampling tables written", flush=True)
    pt, qt, pc, qc, a = od_table_gen()
    np.savez(os.path.join(TABLE_DIR, "od_tables.npz"), photon_prob_rate_total=pt,
             pair_prob_rate_total=qt, integral_photon_prob_along_delta=pc,
             integral_pair_prob_along_delta=qc, chi_N=a["chi_N"],
             log_chi_range=np.asarray(a["log_chi_range"]), log_chi_delta=a["log_chi_delta"],
             delta_N=a["delta_N"], log_delta_range=np.asarray(a["log_delta_range"]),
             log_delta_delta=a["log_delta_delta"])
    print("optical-depth tables written", flush=True)


if __name__ == "__main__":
    main()
