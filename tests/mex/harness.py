"""TEST INFRASTRUCTURE: call the MEX gateways through the fake libmx."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
B = os.path.join(HERE, "_build")
_dp = C.POINTER(C.c_double)


class Mex:
    def __init__(self):
        if not os.path.exists(os.path.join(B, "MCsampleHetro.mexa64")):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        self.mx = C.CDLL(os.path.join(B, "libfakemx.so"), mode=C.RTLD_GLOBAL)
        mx = self.mx
        mx.fake_matrix.restype = C.c_void_p
        mx.fake_matrix.argtypes = [C.c_size_t, C.c_size_t, _dp, _dp]
        mx.fake_struct.restype = C.c_void_p
        mx.fake_set_field.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        mx.fake_rng_reset.argtypes = [C.c_uint64]
        mx.fake_rand.restype = C.c_double
        mx.fake_err_id.restype = C.c_char_p
        mx.fake_err_msg.restype = C.c_char_p
        mx.fake_call.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]
        mx.mxGetPr.restype = _dp
        mx.mxGetPr.argtypes = [C.c_void_p]
        mx.mxGetPi.restype = _dp
        mx.mxGetPi.argtypes = [C.c_void_p]
        mx.mxGetNumberOfElements.restype = C.c_size_t
        mx.mxGetNumberOfElements.argtypes = [C.c_void_p]
        mx.mxGetNumberOfDimensions.restype = C.c_size_t
        mx.mxGetNumberOfDimensions.argtypes = [C.c_void_p]
        mx.mxGetDimensions.restype = C.POINTER(C.c_size_t)
        mx.mxGetDimensions.argtypes = [C.c_void_p]
        mx.mxDestroyArray.argtypes = [C.c_void_p]
        self.gw = {}
        for name in ("MCcov", "MCfield", "MCfieldFourierQuad", "MCfieldOnWave", "MCfieldFourierQuadOnWave",
                     "MCCovHetro", "MCsampleHetro"):
            lib = C.CDLL(os.path.join(B, name + ".mexa64"))
            self.gw[name] = (lib, C.cast(lib.mexFunction, C.c_void_p))

    def arr(self, x):
        if isinstance(x, dict):
            s = self.mx.fake_struct()
            for k, v in x.items():
                self.mx.fake_set_field(s, k.encode(), self.arr(v))
            return s
        a = np.asarray(x)
        if a.ndim == 0:
            a = a.reshape(1, 1)
        if a.ndim == 1:
            a = a.reshape(1, -1)
        re = np.ascontiguousarray(np.real(a).astype(np.float64).reshape(-1, order="F"))
        im = np.ascontiguousarray(np.imag(a).astype(np.float64).reshape(-1, order="F")) if np.iscomplexobj(a) else None
        return self.mx.fake_matrix(a.shape[0], a.shape[1], re.ctypes.data_as(_dp),
                                   im.ctypes.data_as(_dp) if im is not None else None)

    def out(self, p):
        nd = self.mx.mxGetNumberOfDimensions(p)
        dims = [self.mx.mxGetDimensions(p)[i] for i in range(nd)]
        n = self.mx.mxGetNumberOfElements(p)
        re = np.ctypeslib.as_array(self.mx.mxGetPr(p), (n,)).copy()
        pi = self.mx.mxGetPi(p)
        im = np.ctypeslib.as_array(pi, (n,)).copy() if pi else 0
        return (re + 1j * im).reshape(dims, order="F")

    def call(self, name, nlhs, *args):
        """Returns (outputs, None) or (None, (error_id, message))."""
        prhs = (C.c_void_p * len(args))(*[self.arr(a) for a in args])
        plhs = (C.c_void_p * max(nlhs, 1))()
        rc = self.mx.fake_call(self.gw[name][1], nlhs, plhs, len(args), prhs)
        if rc:
            return None, (self.mx.fake_err_id().decode(), self.mx.fake_err_msg().decode())
        return [self.out(plhs[i]) for i in range(nlhs)], None

    def seed_for_next_call(self, rng_seed):
        """The Philox key a gateway derives from the fake MATLAB stream."""
        self.mx.fake_rng_reset(rng_seed)
        u0, u1 = self.mx.fake_rand(), self.mx.fake_rand()
        self.mx.fake_rng_reset(rng_seed)
        return int(u0 * 4294967296.0) | (int(u1 * 4294967296.0) << 32)
