"""Preprocess exported Measurement Set columns into the FIP-TOI input files.

    python -m fastimagingpipe_amd.ingest [--output-prefix P] [--combined-output F]
        [--group-by-time] [--snapshot-workers N] table.npz [table.npz ...]

The options are those of the reference preprocessor (SVDPreprocessMain.cpp:
293-335).  Each .npz holds the columns of one Measurement Set as an external
converter exports them (python-casacore ``getcol``): TIME [nrows], UVW
[nrows, 3], DATA [nrows, nchan, npol] complex, FLAG [nrows, nchan, npol],
optionally WEIGHT_SPECTRUM [nrows, nchan, npol], and CHAN_FREQ (the
SPECTRAL_WINDOW rows, [nspw, nchan] or [nchan]).  Every table must carry the
same CHAN_FREQ and DATA shape.  Without --group-by-time each table is one
snapshot; with it, each distinct TIME is.  Writes <prefix>Bin.fits, Vin.fits,
Visreal.fits, Visimag.fits and the combined file (default <prefix>input.fits).
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

from . import FipError, ingest_table


def load_tables(paths):
    """Concatenated columns of the tables with the table index of each row."""
    cols = {"TIME": [], "UVW": [], "DATA": [], "FLAG": [], "WEIGHT_SPECTRUM": [], "file": []}
    freqs = None
    shape = None
    for k, path in enumerate(paths):
        with np.load(path, allow_pickle=False) as z:
            f = np.asarray(z["CHAN_FREQ"], np.float64).reshape(-1)
            d = np.asarray(z["DATA"])
            if freqs is None:
                freqs, shape = f, d.shape[1:]
            elif not np.array_equal(f, freqs) or d.shape[1:] != shape:
                raise ValueError(f"{path}: CHAN_FREQ or DATA shape differs from the first table")
            n = d.shape[0]
            cols["TIME"].append(np.asarray(z["TIME"], np.float64).reshape(n))
            cols["UVW"].append(np.asarray(z["UVW"], np.float64).reshape(n, 3))
            cols["DATA"].append(d)
            cols["FLAG"].append(np.asarray(z["FLAG"]).reshape(d.shape))
            cols["WEIGHT_SPECTRUM"].append(np.asarray(z["WEIGHT_SPECTRUM"], np.float32).reshape(d.shape)
                                           if "WEIGHT_SPECTRUM" in z.files else None)
            cols["file"].append(np.full(n, k, np.int32))
    ws = cols["WEIGHT_SPECTRUM"]
    if all(w is None for w in ws):
        weight = None
    else:
        weight = np.concatenate([w if w is not None else np.ones(d.shape, np.float32)
                                 for w, d in zip(ws, cols["DATA"])])
    return dict(time=np.concatenate(cols["TIME"]), uvw=np.concatenate(cols["UVW"]),
                data=np.concatenate(cols["DATA"]), flag=np.concatenate(cols["FLAG"]), weight=weight,
                file_index=np.concatenate(cols["file"]), freqs=freqs)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("tables", nargs="+")
    ap.add_argument("--output-prefix", default="")
    ap.add_argument("--combined-output", default=None)
    ap.add_argument("--group-by-time", action="store_true")
    ap.add_argument("--snapshot-workers", type=int, default=None)
    a = ap.parse_args(argv)
    if a.snapshot_workers is not None and a.snapshot_workers < 1:
        ap.error("--snapshot-workers must be greater than zero")
    try:
        t = load_tables(a.tables)
        ns, np_ = ingest_table(t["uvw"], t["data"], t["flag"], t["freqs"], time=t["time"],
                               file_index=t["file_index"], weight=t["weight"],
                               group_by_time=a.group_by_time, workers=a.snapshot_workers or 0,
                               prefix=a.output_prefix, combined=a.combined_output)
    except (FipError, ValueError, KeyError, OSError) as e:
        print(f"Preprocessing failed: {e}", file=sys.stderr)
        return 1
    print(f"Wrote {ns} snapshot(s), {np_} padded samples per snapshot; combined FITS input: "
          f"{a.combined_output or a.output_prefix + 'input.fits'}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
