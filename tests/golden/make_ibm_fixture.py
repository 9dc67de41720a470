"""Converts the reference's data fixture data/Umtx.mat (the 4th IBM Challenge
4-qubit unitary, MATLAB v5) into ibm_challenge_umtx.npy beside this script.
scipy.io.loadmat reads the MATLAB format without executing anything."""
import os
import sys

import numpy as np
from scipy.io import loadmat

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data/Umtx.mat"
U = np.ascontiguousarray(loadmat(src)["Umtx"], dtype=np.complex128)
np.save(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ibm_challenge_umtx.npy"), U)
print(U.shape)
