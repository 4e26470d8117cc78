"""Write the golden fixtures for the WarpPerRow codec by running the reference generator.

Runs ``matvansWarpPerRow/create_matrices_WarpPerRow.py`` (reference tree, read-only)
as a subprocess in a scratch directory, then copies its ``.bin`` outputs here:

  compressed_matrices_512_5bit_count2.bin   -w 512  -n 2 -e 5   (one granule per row)
  raw-matrices_512.bin.gz                   -w 512  -n 2 -e 5 -u (gzip of the raw file)
  compressed_matrices_1024_5bit_count1.bin  -w 1024 -n 1 -e 5   (two granules per row,
                                                                 two 512-column groups)

The generator is a module-level script (argparse + multiprocessing at import), so it
can only be run, not imported.  Only functions that no closed form pins are taken
from it: the byte-exact entropy model / ppf / interleave (numpy argsort tie-breaking,
``np.round`` half-even, ``np.empty`` padding) and the float64-chain ``result_hash``.

The generator only works when ``max - min <= 127`` or ``G + 1 == 256``: it indexes
``padded_pmf[value - min_value + 1]`` with an int8 ``value``, which wraps under
numpy 2 promotion (create_matrices_WarpPerRow.py:180-181).  ``-e 5`` on these sizes
hits both +-127, so the wrap lands on the right slot by accident; ``-e 8`` asserts.

Usage: python tests/golden/make_fixtures.py [--ref /root/reference]
"""
import argparse
import gzip
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

RUNS = [
    (["-w", "512", "-n", "2", "-e", "5", "-u"],
     {"compressed_matrices_512_5bit_count2.bin": "compressed_matrices_512_5bit_count2.bin",
      "raw-matrices_512.bin": "raw-matrices_512.bin.gz"}),
    (["-w", "1024", "-n", "1", "-e", "5"],
     {"compressed_matrices_1024_5bit_count1.bin": "compressed_matrices_1024_5bit_count1.bin"}),
]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    script = os.path.join(args.ref, "matvansWarpPerRow", "create_matrices_WarpPerRow.py")
    manifest = {}
    for argv, outputs in RUNS:
        with tempfile.TemporaryDirectory() as tmp:
            cmd = [sys.executable, script] + argv
            res = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True)
            if res.returncode != 0:
                sys.stderr.write(res.stdout[-2000:] + res.stderr[-4000:])
                raise SystemExit(f"reference generator failed: {' '.join(argv)}")
            for src, dst in outputs.items():
                data = open(os.path.join(tmp, src), "rb").read()
                manifest[dst] = {"args": argv, "sha256_raw": hashlib.sha256(data).hexdigest(),
                                 "bytes_raw": len(data)}
                out = os.path.join(HERE, dst)
                if dst.endswith(".gz"):
                    with gzip.GzipFile(out, "wb", compresslevel=9, mtime=0) as f:
                        f.write(data)
                else:
                    shutil.copyfile(os.path.join(tmp, src), out)
    import numpy
    manifest["_generator"] = {"numpy": numpy.__version__, "python": sys.version.split()[0]}
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print(json.dumps(manifest, indent=1))


if __name__ == "__main__":
    main()
