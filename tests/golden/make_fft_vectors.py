"""Writes tests/golden/fft_vectors.npz: numpy (pocketfft, FP64) transforms of
seeded inputs, the golden vectors the oracle and the HIP FFT are checked on."""
import os

import numpy as np

rng = np.random.Generator(np.random.Philox(20261203))
x1 = rng.standard_normal(2048) + 1j * rng.standard_normal(2048)
x2 = np.asfortranarray(rng.standard_normal((64, 128)) + 1j * rng.standard_normal((64, 128)))
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_vectors.npz"),
                    x1d=x1, X1d=np.fft.fft(x1), x2d=x2, X2d=np.fft.fft2(x2))
