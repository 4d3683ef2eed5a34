"""Writes the golden FFT vectors from the reference's own Ooura FFT.

Runs oracle/_ref/libooura_ref.so (fftsg3d.c + fftsg.c from the reference,
built by oracle/Makefile) through the call sequence of the reference's
A2a_k / A_k2a (fft_wrapper.h:681-723) on seeded inputs, and stores inputs
and outputs.  Run in the build container:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.oracle import OouraRef  # noqa: E402


def main():
    ref = OouraRef()
    for shape in [(16, 16, 16), (32, 16, 32)]:
        nx, ny, nz = shape
        rng = np.random.default_rng(1000 + nx + ny + nz)
        a = np.zeros((nx, ny, nz + 2))
        a[:, :, :nz] = rng.standard_normal((nx, ny, nz))
        spec = ref.a2a_k(a)
        # inverse input: a Hermitian-consistent spectrum (the forward output
        # truncated by the 2/3 rule so no Nyquist mode survives)
        trn = [(n + 2) // 3 for n in shape]
        kx = np.fft.fftfreq(nx, 1.0 / nx).astype(int)
        ky = np.fft.fftfreq(ny, 1.0 / ny).astype(int)
        kz = np.arange(nz // 2 + 1)
        keep = (np.abs(kx)[:, None, None] < trn[0]) & (np.abs(ky)[None, :, None] < trn[1]) & \
               (kz[None, None, :] < trn[2])
        sc = spec.view(np.complex128) * keep
        spec_in = sc.view(np.float64).copy()
        back = ref.a_k2a(spec_in)
        np.savez_compressed(os.path.join(HERE, f"ooura_fft_{nx}x{ny}x{nz}.npz"), real_in=a[:, :, :nz],
                            spec_out=spec, spec_in=spec_in, real_out=back[:, :, :nz])
        print("wrote", shape)


if __name__ == "__main__":
    main()
