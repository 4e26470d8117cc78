"""Write the golden fixtures by running the reference generators.

Runs ``matvansWarpPerRow/create_matrices_WarpPerRow.py`` and ``buffered/encoder.py``
(reference tree, read-only) as subprocesses in a scratch directory, then copies their
outputs here.  WarpPerRow (.bin v0):

  compressed_matrices_512_5bit_count2.bin   -w 512  -n 2 -e 5   (one granule per row)
  raw-matrices_512.bin.gz                   -w 512  -n 2 -e 5 -u (gzip of the raw file)
  compressed_matrices_1024_5bit_count1.bin  -w 1024 -n 1 -e 5   (two granules per row,
                                                                 two 512-column groups)

Family B (quad rANS, safetensors container), ``buffered/encoder.py``:

  b_rows_256.safetensors      --dim 256 --count 2 --mode ans-warp-rows
  b_stripes_256.safetensors   --dim 256 --count 2 --mode ans-warp-stripes
  b_raw_256.safetensors       --dim 256 --count 2 --mode uncompressed
  b_rows_384_e5.safetensors   --dim 384 --count 1 --target_entropy 5 --mode ans-warp-rows
  b_stripes_384_e5.safetensors  (same, ans-warp-stripes)
  b_raw_384_e5.safetensors      (same, uncompressed)

encoder.py cannot encode more than 128 distinct values (its 7-bit model then gives some
symbol probability 0 and the coder divides by zero, encoder.py:156-166,221), so
``--target_entropy 6`` fails at these sizes.  The WarpPerRow generator leaves padding
slots as ``np.empty`` garbage, so rerunning it changes bytes outside the streams:
``--only b`` regenerates the family-B files alone.

From encoder.py only what no closed form pins is taken: the byte-exact 7-bit model
(float32 rounding, ``np.argmax`` adjustment order), the coders' interleave and the
per-matrix ``hashes`` (``np.round`` of int32 x float32, promoted to float64 by numpy).

The WarpPerRow generator is a module-level script (argparse + multiprocessing at
import), so it can only be run, not imported.  Only functions that no closed form pins are taken
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

B_RUNS = [
    (["--dim", "256", "--count", "2"], "256"),
    (["--dim", "384", "--count", "1", "--target_entropy", "5"], "384_e5"),
]
B_MODES = [("ans-warp-rows", "b_rows"), ("ans-warp-stripes", "b_stripes"),
           ("uncompressed", "b_raw")]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", choices=["v0", "b"], default=None)
    args = ap.parse_args()
    script = os.path.join(args.ref, "matvansWarpPerRow", "create_matrices_WarpPerRow.py")
    mpath = os.path.join(HERE, "MANIFEST.json")
    manifest = json.load(open(mpath)) if os.path.exists(mpath) else {}
    for argv, outputs in (RUNS if args.only != "b" else []):
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
    bscript = os.path.join(args.ref, "buffered", "encoder.py")
    for argv, tag in (B_RUNS if args.only != "v0" else []):
        for mode, prefix in B_MODES:
            name = f"{prefix}_{tag}.safetensors"
            with tempfile.TemporaryDirectory() as tmp:
                full = argv + ["--mode", mode, "-o", name]
                res = subprocess.run([sys.executable, bscript] + full, cwd=tmp,
                                     capture_output=True, text=True)
                if res.returncode != 0:
                    sys.stderr.write(res.stdout[-2000:] + res.stderr[-4000:])
                    raise SystemExit(f"reference encoder failed: {' '.join(full)}")
                data = open(os.path.join(tmp, name), "rb").read()
                manifest[name] = {"args": full, "sha256_raw": hashlib.sha256(data).hexdigest(),
                                  "bytes_raw": len(data)}
                shutil.copyfile(os.path.join(tmp, name), os.path.join(HERE, name))
    import numpy
    manifest["_generator"] = {"numpy": numpy.__version__, "python": sys.version.split()[0]}
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print(json.dumps(manifest, indent=1))


if __name__ == "__main__":
    main()
