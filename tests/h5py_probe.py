"""Independent reader/writer of run files with h5py, for the run-file tests.

The tests run this script with an interpreter that has h5py (the package
itself does not need it):

  h5py_probe.py describe FILE
      prints JSON: for each dataset its dtype, shape, maxshape, chunks and the
      float64 sum and first values of its data;
  h5py_probe.py write FILE NPZ RECORDS D0 D1 D2 D3 D4 D5
      writes a run file with the h5py calls of RunData.create_file and
      RunData.save_data (main/data_management.py:13-52) -- create_dataset with
      chunks=True, maxshape (None, ...), dtype 'f', "time"/"energy" from
      data=[0], then resize + write of one record per save -- and saves the
      records it wrote (pdf, potential, density, time, energy) to NPZ.
"""
import json
import sys

import h5py
import numpy as np


def describe(path):
    out = {}
    with h5py.File(path, "r") as f:
        for name in f:
            d = f[name]
            a = d[()]
            out[name] = {"dtype": str(d.dtype), "shape": list(d.shape),
                         "maxshape": [-1 if m is None else m for m in d.maxshape],
                         "chunks": list(d.chunks) if d.chunks else None,
                         "sum": float(np.asarray(a, dtype=np.float64).sum()),
                         "head": np.asarray(a, dtype=np.float64).ravel()[:4].tolist()}
    print(json.dumps(out))


def write(path, npz, records, shape):
    rows, cols = shape[0] - 2, shape[1]
    size = int(np.prod(shape))
    pdfs, pots, dens, times, energies = [], [], [], [], []
    for r in range(records):
        k = np.arange(size, dtype=np.float64)
        pdfs.append((np.sin(0.37 * k + r) * 10.0 ** (-(k % 5))).reshape(shape).astype(np.float32))
        j = np.arange(rows * cols, dtype=np.float64).reshape(rows, cols)
        pots.append(np.cos(0.5 * j + r).astype(np.float32))
        dens.append((1.0 + 1e-3 * np.sin(j - r)).astype(np.float32))
        times.append(float(r))          # integral times survive the int64 datasets
        energies.append(0.25 * r)
    with h5py.File(path, "w") as f:
        f.create_dataset("pdf", data=np.array([pdfs[0]]), chunks=True, maxshape=(None,) + tuple(shape), dtype="f")
        f.create_dataset("potential", data=np.array([pots[0]]), chunks=True, maxshape=(None, shape[0], shape[1]),
                         dtype="f")
        f.create_dataset("density", data=np.array([dens[0]]), chunks=True, maxshape=(None, shape[0], shape[1]),
                         dtype="f")
        f.create_dataset("time", data=[0], chunks=True, maxshape=(None,))
        f.create_dataset("energy", data=[0], chunks=True, maxshape=(None,))
    for r in range(1, records):
        with h5py.File(path, "a") as f:
            for name in ("pdf", "potential", "density", "time", "energy"):
                f[name].resize((f[name].shape[0] + 1), axis=0)
            f["pdf"][-1] = pdfs[r]
            f["potential"][-1] = pots[r]
            f["density"][-1] = dens[r]
            f["time"][-1] = times[r]
            f["energy"][-1] = energies[r]
    np.savez(npz, pdf=np.array(pdfs), potential=np.array(pots), density=np.array(dens),
             time=np.array(times), energy=np.array(energies))


if __name__ == "__main__":
    if sys.argv[1] == "describe":
        describe(sys.argv[2])
    else:
        write(sys.argv[2], sys.argv[3], int(sys.argv[4]), [int(x) for x in sys.argv[5:11]])
