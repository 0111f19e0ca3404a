"""Pack AMDAT's q-vector shell files into compact test fixtures.

Reads src/qvectors/qvectors.zip from the reference checkout (files
qvectors3d/qvector.001 .. qvector.300: integer triples written with %e,
CRLF line ends; src/wave_vectors.cpp:430-484 reads them) and writes
tests/golden/qvectors3d.npz with

    n    int8  [total, 3]  the integer vectors, shells concatenated in
                           file order, vectors in in-file order
    off  int32 [301]       shell b (file qvector.%03d % b) holds rows
                           off[b-1] .. off[b]-1

Also writes, from the same archive,

    qvectors2d.npz  n int8 [total, 3], off int32 [286]: files
                    qvectors2d/qvector.000 .. .284 (tab-separated integer
                    triples; src/wave_vectors.cpp:266-336 reads them);
                    shell b holds rows off[b] .. off[b+1]-1
    qvectors1d.npz  n2 int16 [300, 3]: twice the one triple of each file
                    qvectors1d/qvector.001 .. .300 (half-integers bin/2;
                    src/wave_vectors.cpp:343-424 reads them)

The data is AMDAT's (GPL-3.0, (C) David S. Simmons); it is kept here only
as a test fixture so the shell loader and every parity test can be
checked against the exact vectors the reference uses.

Run once in a container that has the reference checkout:
    python tests/golden/make_qvector_fixture.py /root/reference
"""
import io
import os
import sys
import zipfile

import numpy as np


def main(ref_root):
    zpath = os.path.join(ref_root, "src", "qvectors", "qvectors.zip")
    rows, off = [], [0]
    with zipfile.ZipFile(zpath) as zf:
        for b in range(1, 301):
            text = zf.read("qvectors3d/qvector.%03d" % b).decode("ascii")
            vals = [float(t) for t in text.split()]
            assert len(vals) % 3 == 0
            trip = np.array(vals, dtype=np.float64).reshape(-1, 3)
            assert np.all(trip == np.round(trip))
            rows.append(trip.astype(np.int8))
            off.append(off[-1] + trip.shape[0])
    n = np.concatenate(rows, axis=0)
    here = os.path.dirname(os.path.abspath(__file__))
    out = os.path.join(here, "qvectors3d.npz")
    np.savez_compressed(out, n=n, off=np.array(off, dtype=np.int32))
    print("wrote", out, n.shape)
    with zipfile.ZipFile(zpath) as zf:
        rows, off = [], [0]
        for b in range(0, 285):
            vals = [float(t) for t in zf.read("qvectors2d/qvector.%03d" % b).decode("ascii").split()]
            assert len(vals) % 3 == 0
            trip = np.array(vals, dtype=np.float64).reshape(-1, 3)
            assert np.all(trip == np.round(trip)) and np.all(np.abs(trip) <= 127)
            rows.append(trip.astype(np.int8))
            off.append(off[-1] + trip.shape[0])
        n2d = np.concatenate(rows, axis=0)
        out = os.path.join(here, "qvectors2d.npz")
        np.savez_compressed(out, n=n2d, off=np.array(off, dtype=np.int32))
        print("wrote", out, n2d.shape)
        one = []
        for b in range(1, 301):
            vals = [float(t) for t in zf.read("qvectors1d/qvector.%03d" % b).decode("ascii").split()]
            assert len(vals) == 3
            two = 2.0 * np.array(vals)
            assert np.all(two == np.round(two))
            one.append(two.astype(np.int16))
        out = os.path.join(here, "qvectors1d.npz")
        np.savez_compressed(out, n2=np.array(one))
        print("wrote", out, len(one))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
