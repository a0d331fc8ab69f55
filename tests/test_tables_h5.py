"""Reading the reference's HDF5 table files (h5py is not a dependency: the
format tests run where it is installed and are skipped otherwise).

The round trip writes the shipped tables in the reference's layout (dataset
attributes for the RS file, file attributes for the OD file, as
qed/rejection_sampling_tables.py:96-112 and qed/optical_depth_tables.py:206-230
write them) and reads them back bit for bit.  use_tables swaps the tables the
next simulation uploads.
"""
import numpy as np
import pytest


def test_use_tables_swaps_the_host_tables():
    from sfparticles_amd import qed
    from sfparticles_amd.qed import tables as T
    rs0, od0 = T.load_rs_tables(), T.load_od_tables()
    rs1 = T.RSTables(rs0.int_Ai * 1.0, rs0.Aip * 1.0, rs0.z_N, rs0.z_min, rs0.z_max, rs0.z_delta,
                     rs0.Ai)
    T.use_tables(rs=rs1)
    try:
        assert T.load_rs_tables() is rs1 and T.load_od_tables() is od0
        assert qed._tables_handle is None
    finally:
        T.use_tables(rs=rs0)
    assert T.load_rs_tables() is rs0


def test_reader_needs_h5py_or_says_so(tmp_path):
    from sfparticles_amd.qed import tables as T
    try:
        import h5py  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError, match="h5py"):
            T.read_rs_tables_h5(str(tmp_path / "none.h5"))


def test_h5_round_trip(tmp_path):
    h5py = pytest.importorskip("h5py")
    from sfparticles_amd.qed import tables as T
    rs, od = T.load_rs_tables(), T.load_od_tables()
    p_rs, p_od = tmp_path / "rejection_sampling_tables.h5", tmp_path / "optical_depth_tables.h5"
    with h5py.File(p_rs, "w") as f:
        for name, arr in (("int_Ai", rs.int_Ai), ("Aip", rs.Aip)):
            d = f.create_dataset(name, data=arr)
            d.attrs["z_N"] = rs.z_N
            d.attrs["z_range"] = (rs.z_min, rs.z_max)
            d.attrs["z_delta"] = rs.z_delta
    with h5py.File(p_od, "w") as f:
        f.create_dataset("photon_prob_rate_total", data=od.photon_total)
        f.create_dataset("pair_prob_rate_total", data=od.pair_total)
        f.create_dataset("integral_photon_prob_along_delta", data=od.photon_cdf)
        f.create_dataset("integral_pair_prob_along_delta", data=od.pair_cdf)
        f.attrs["chi_N"] = od.chi_N
        f.attrs["log_chi_range"] = (od.log_chi_min, od.log_chi_max)
        f.attrs["log_chi_delta"] = od.log_chi_delta
        f.attrs["delta_N"] = od.delta_N
        f.attrs["log_delta_range"] = (od.log_delta_min, od.log_delta_max)
        f.attrs["log_delta_delta"] = od.log_delta_delta
    rs2, od2 = T.read_rs_tables_h5(str(p_rs)), T.read_od_tables_h5(str(p_od))
    assert np.array_equal(rs2.int_Ai, rs.int_Ai) and np.array_equal(rs2.Aip, rs.Aip)
    assert (rs2.z_N, rs2.z_min, rs2.z_max, rs2.z_delta) == (rs.z_N, rs.z_min, rs.z_max, rs.z_delta)
    # Ai is regenerated on the grid (scipy's airy: equal up to its version's ulps)
    assert np.allclose(rs2.Ai, rs.Ai, rtol=1e-12, atol=0)
    for a in ("photon_total", "pair_total", "photon_cdf", "pair_cdf", "chi_grid", "delta_grid"):
        assert np.array_equal(getattr(od2, a), getattr(od, a)), a
