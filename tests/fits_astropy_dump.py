"""Read FITS files with astropy.io.fits, an independent FITS implementation,
and save what it sees for the output-format tests to compare with.

Run by tests/test_output_formats.py under the image's Python 3.9 (which
ships astropy 4.3.1), not under the interpreter that runs the tests:

    python3.9 fits_astropy_dump.py OUT.npz FILE [FILE ...]

For every file it first runs astropy's standard-conformance check
(verify('exception')), then stores each HDU's header as JSON and its data
(image arrays; binary-table columns by name) in OUT.npz.
"""
import json
import sys
import warnings

import numpy as np

# astropy 4.3 imports two helpers that newer numpy releases dropped; give
# them back before astropy's table module is loaded.
if not hasattr(np, "asscalar"):
    np.asscalar = lambda a: a.item()
if not hasattr(np, "alen"):
    np.alen = len

warnings.simplefilter("ignore")
from astropy.io import fits  # noqa: E402


def main(out, paths):
    arrays = {}
    for i, path in enumerate(paths):
        with fits.open(path) as hdul:
            hdul.verify("exception")
            for j, hdu in enumerate(hdul):
                head = {k: v for k, v in hdu.header.items() if k not in ("COMMENT", "HISTORY", "")}
                arrays[f"f{i}_h{j}_header"] = np.frombuffer(json.dumps(head).encode(), np.uint8)
                if isinstance(hdu, fits.BinTableHDU):
                    for name in hdu.columns.names:
                        arrays[f"f{i}_h{j}_col_{name}"] = np.asarray(hdu.data[name])
                elif hdu.data is not None:
                    arrays[f"f{i}_h{j}_data"] = np.asarray(hdu.data)
            arrays[f"f{i}_nhdu"] = np.array(len(hdul))
    np.savez(out, **arrays)


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2:])
