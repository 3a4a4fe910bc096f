"""Writes tests/golden/encode_formats.npz: lightmap texels written into the buffer formats the
reference's buffer interface supports per texel (oracle/_ref/encode_ref: the writeback with
the reference's own sRGB colour space, RR_FLOAT2BYTE and RRVec3::avg).

Run in a container that has /root/reference (`make -C oracle ref integration`):

    python tests/golden/make_encode_golden.py

Inputs: 2051 seeded RGBA texels (a tail that is not a multiple of 4): values in [0, 2),
tiny values, values on and next to every byte step k/256 and to the sRGB steps
(k/256)^(1/0.45), negative values, values whose x*256 leaves the int range, inf and NaN,
and background texels (alpha 0 or negative); background colour (0.2, 0.5, 0.8, 0.7).
Outputs: one byte array per (format, scaled target, bent-normal layer)."""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(ROOT, "oracle", "_ref", "encode_ref")
FORMATS = (0, 1, 2, 3, 4, 9, 10)
BACKGROUND = np.float32([0.2, 0.5, 0.8, 0.7])


def inputs(n=2051, seed=12):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 2, (n, 4)).astype(np.float32)
    k = np.arange(256, dtype=np.float64)
    steps = np.concatenate([k / 256, (k / 256) ** (1 / 0.45)]).astype(np.float32)
    near = np.concatenate([steps, np.nextafter(steps, np.float32(-1)), np.nextafter(steps, np.float32(3))])
    flat = x[:, :3].reshape(-1)
    flat[:len(near)] = near
    flat[len(near):len(near) + 40] = rng.uniform(0, 1e-6, 40)
    flat[len(near) + 40:len(near) + 80] = -rng.uniform(0, 0.5, 40)
    flat[len(near) + 80:len(near) + 90] = [1e7, 9e6, 8.4e6, 8.3e6, -1e8, 3e9, np.inf, -np.inf, np.nan, 255.9 / 256]
    x[:, :3] = flat.reshape(-1, 3)
    x[:, 3] = rng.uniform(0.01, 1, n)
    x[::7, 3] = 0
    x[3::11, 3] = -1
    x[5::13, 3] = 0.001
    return x


def encode(rgba, fmt, scaled, bent, tmp):
    fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    with open(fin, "wb") as f:
        f.write(np.uint32(len(rgba)).tobytes())
        f.write(np.ascontiguousarray(rgba, np.float32).tobytes())
        f.write(BACKGROUND.tobytes())
    subprocess.run([BIN, fin, fout, str(fmt), str(scaled), str(bent)], check=True)
    return np.fromfile(fout, np.uint8)


def main():
    import tempfile
    if not os.path.exists(BIN):
        sys.exit("build oracle/_ref/encode_ref first: make -C oracle ref integration")
    x = inputs()
    out = dict(rgba=x, background=BACKGROUND)
    with tempfile.TemporaryDirectory() as tmp:
        for fmt in FORMATS:
            for scaled in (0, 1):
                for bent in (0, 1):
                    out[f"f{fmt}_s{scaled}_b{bent}"] = encode(x, fmt, scaled, bent, tmp)
    np.savez_compressed(os.path.join(HERE, "encode_formats.npz"), **out)
    print("wrote", os.path.join(HERE, "encode_formats.npz"))


if __name__ == "__main__":
    main()
