"""Pack AMDAT's 3D q-vector shell files into a compact test fixture.

Reads src/qvectors/qvectors.zip from the reference checkout (files
qvectors3d/qvector.001 .. qvector.300: integer triples written with %e,
CRLF line ends; src/wave_vectors.cpp:430-484 reads them) and writes
tests/golden/qvectors3d.npz with

    n    int8  [total, 3]  the integer vectors, shells concatenated in
                           file order, vectors in in-file order
    off  int32 [301]       shell b (file qvector.%03d % b) holds rows
                           off[b-1] .. off[b]-1

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
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "qvectors3d.npz")
    np.savez_compressed(out, n=n, off=np.array(off, dtype=np.int32))
    print("wrote", out, n.shape)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
