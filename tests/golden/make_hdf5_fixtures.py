"""Writes the HDF5 line-list fixtures of tests/test_hdf5_lists.py in the layout of the
reference's converters (python/vald2hdf.py:763-842, python/mol2hdf.py): a /Lines group with
one dataset per column -- W f8 contiguous, the other columns chunked with the shuffle and
gzip (level 4) filters -- and the group attributes NL, NAME, WL_VAC, HFLAG, WMIN, WMAX
written as that converter writes them (strings fixed-length and NUL-terminated).  The
columns come from the package's seeded generators, so a test regenerates them.

Needs h5py; in this image it lives in the conda Python 3.9:

    /opt/conda/bin/python3.9 tests/golden/make_hdf5_fixtures.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from fastspec_amd import workloads as W  # noqa: E402

OUT = Path(__file__).resolve().parent
# (file, kind, lines, wavelength range, seed, rows per chunk, WL_VAC); chunk 0: the
# converter's own choice, min(n, 262144)
FIXTURES = [
    ("atomic_ref.h5", "atomic", 3000, (5000.0, 5100.0), 901, 0, 0),
    ("atomic_chunked_vac.h5", "atomic", 20000, (6000.0, 6200.0), 902, 1024, 1),
    ("feh_chunked.h5", "molecular", 20000, (9000.0, 9100.0), 903, 1000, 0),
    ("zeeman_chunked.h5", "zeeman", 600, (6000.0, 6010.0), 904, 128, 0),
]


def columns(kind, n, rng, seed):
    if kind == "molecular":
        cols = W.diatomic_columns(n, *rng, seed)
        # wavelengths as sorted uniform draws: exp() of the log-uniform draws is not the
        # same to the last bit in every NumPy version, and the tests regenerate the columns
        # with the NumPy they run under
        cols["W"] = np.sort(np.random.Generator(np.random.PCG64(seed)).uniform(*rng, n))
        return cols
    cols = W.atomic_columns(n, *rng, seed)
    if kind == "zeeman":
        cols.update(W.zeeman_level_columns(n, seed + 1))
    return cols


def write(path, kind, cols, chunk, wl_vac):
    from fastspec_amd import linelist as LL
    name = "FeH" if kind == "molecular" else ("atom Zeeman" if kind == "zeeman" else "atom")
    LL.write_hdf5(path, cols, name=name, wl_vac=wl_vac, chunk=chunk or (1 << 18), mol_id="0126.001056")


def main():
    for fn, kind, n, rng, seed, chunk, vac in FIXTURES:
        write(OUT / fn, kind, columns(kind, n, rng, seed), chunk, vac)
        print(fn, (OUT / fn).stat().st_size)


if __name__ == "__main__":
    main()
