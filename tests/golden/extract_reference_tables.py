"""Extract the data the reference ships for the particle step into .npz fixtures.

Run in the development container only (it reads the reference checkout):

    /opt/conda/bin/python3.9 tests/golden/extract_reference_tables.py /root/reference

What it writes (data only -- no reference code is copied or executed):
  rs_tables_reference.npz   datasets `int_Ai`, `Aip` and the attributes `z_N`,
                            `z_range`, `z_delta` of
                            sfparticles/qed/rejection_sampling_tables.h5, read with
                            h5py (a plain data loader; HDF5 executes nothing).
  photon_total_kat.npy      the 256-value known-answer vector of the photon total
                            rate on chi = logspace(-3, 2, 256) in units of
                            (sqrt(3)/2pi) alpha m c^2/hbar, parsed as numbers from
                            tests/test_tables.py:11 with ast.literal_eval.
"""
import ast
import os
import sys

import numpy as np


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    import h5py

    path = os.path.join(ref_root, "sfparticles", "qed", "rejection_sampling_tables.h5")
    with h5py.File(path, "r") as f:
        int_Ai = np.asarray(f["int_Ai"][()], dtype=np.float64)
        Aip = np.asarray(f["Aip"][()], dtype=np.float64)
        attrs = {k: np.asarray(v) for k, v in f["Aip"].attrs.items()}
        attrs_int = {k: np.asarray(v) for k, v in f["int_Ai"].attrs.items()}
    np.savez(
        os.path.join(here, "rs_tables_reference.npz"),
        int_Ai=int_Ai, Aip=Aip,
        z_N=attrs["z_N"], z_range=attrs["z_range"], z_delta=attrs["z_delta"],
        int_Ai_z_N=attrs_int["z_N"], int_Ai_z_range=attrs_int["z_range"],
        int_Ai_z_delta=attrs_int["z_delta"],
    )

    src = open(os.path.join(ref_root, "tests", "test_tables.py")).read()
    line = [l for l in src.splitlines() if l.startswith("prob_rate_table = [")][0]
    values = ast.literal_eval(line.split("=", 1)[1].strip())
    np.save(os.path.join(here, "photon_total_kat.npy"), np.asarray(values, dtype=np.float64))
    print("int_Ai", int_Ai.shape, "Aip", Aip.shape, "kat", len(values))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
