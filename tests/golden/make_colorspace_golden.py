"""Writes tests/golden/colorspace_srgb.npz: the reference's sRGB colour space
(RRColorSpace::create_sRGB(0.45), compiled from its sources into oracle/_ref/colorspace_ref)
applied to texel values as RRBufferInMemory::getElement applies it: getLinear of every byte
value (the byte table of scaled RGB and RGBA textures) and toLinear of sample floats.

Run in a container that has /root/reference (`make -C oracle integration`):

    python tests/golden/make_colorspace_golden.py

Inputs: gamma 0.45; every byte value 0..255; 4096 seeded floats in [0, 2), every byte step
k/255 and its float neighbours, the byte decodes k * RR_BYTE2FLOAT, tiny and negative values.
Outputs: byte_linear[256], x[n], to_linear[n] (float32)."""
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(ROOT, "oracle", "_ref", "colorspace_ref")
GAMMA = 0.45


def inputs(seed=3):
    rng = np.random.default_rng(seed)
    k = np.arange(256, dtype=np.float32)
    steps = k / np.float32(255)
    near = np.concatenate([steps, np.nextafter(steps, np.float32(-1)), np.nextafter(steps, np.float32(3))])
    decodes = k * np.float32(0.003921568627450980392156862745098)
    x = np.concatenate([rng.uniform(0, 2, 4096).astype(np.float32), near, decodes,
                        rng.uniform(0, 1e-6, 64).astype(np.float32), -rng.uniform(0, 0.5, 64).astype(np.float32)])
    return x.astype(np.float32)


def main():
    x = inputs()
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<I", len(x)) + x.tobytes())
        subprocess.run([BIN, fin, fout, str(GAMMA)], check=True)
        o = np.fromfile(fout, np.float32)
    np.savez_compressed(os.path.join(HERE, "colorspace_srgb.npz"), gamma=np.float32(GAMMA), byte_linear=o[:256], x=x,
                        to_linear=o[256:])
    print(f"wrote {len(x)} toLinear samples and the byte table")


if __name__ == "__main__":
    main()
