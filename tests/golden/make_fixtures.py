"""Writes the reference-held fixture as plain arrays.

Source: examples/buildFarFieldCovMatrix/scatteringAmplitude.mat of the
reference repository (MATLAB v5 data: theta 1x3601 on [0, 2pi], complex
farField 1x3601 -- the 2-D mu-diff amplitude used by farFieldCov.m:21-26).
Read with scipy.io.loadmat (a data parser), written as a numpy .npz.
Run once in the container that holds the reference; the GPU box uses the .npz.
"""
import os
import sys

import numpy as np
import scipy.io as sio

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

m = sio.loadmat(os.path.join(REF, "examples", "buildFarFieldCovMatrix", "scatteringAmplitude.mat"))
theta = np.asarray(m["theta"], dtype=np.float64).reshape(-1)
ff = np.asarray(m["farField"], dtype=np.complex128).reshape(-1)
np.savez_compressed(os.path.join(HERE, "scatteringAmplitude.npz"), theta=theta, farField_re=ff.real,
                    farField_im=ff.imag)
print("wrote", theta.size, "samples; theta[-1] == 2*pi:", theta[-1] == 2 * np.pi)
