"""ctypes driver of the MEX gateways built against tests/fake_mex (test infrastructure)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FAKE = os.path.join(HERE, "fake_mex")


def build():
    subprocess.run(["make", "-s", "-C", FAKE], check=True)


def load(which):
    path = os.path.join(FAKE, "build", f"libfakemex_{which}.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    vp = C.c_void_p
    L.fm_double.restype = vp
    L.fm_double.argtypes = [C.POINTER(C.c_double), C.c_size_t, C.POINTER(C.c_size_t)]
    L.fm_empty.restype = vp
    L.fm_string.restype = vp
    L.fm_string.argtypes = [C.c_char_p]
    L.fm_object.restype = vp
    L.fm_object.argtypes = [C.c_char_p, C.c_char_p]
    L.fm_set_prop.argtypes = [vp, C.c_char_p, vp]
    L.fm_call.argtypes = [C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp), C.c_char_p, C.c_size_t]
    L.fm_numel.restype = C.c_size_t
    L.fm_numel.argtypes = [vp]
    L.fm_dim.restype = C.c_size_t
    L.fm_dim.argtypes = [vp, C.c_size_t]
    L.fm_is_single_complex.argtypes = [vp]
    L.fm_get_single.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.fm_is_double.argtypes = [vp]
    L.fm_get_double.argtypes = [vp, C.POINTER(C.c_double)]
    L.fm_single_complex.restype = vp
    L.fm_single_complex.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_size_t, C.c_size_t]
    L.fm_free.argtypes = [vp]
    return Gateway(L)


class Gateway:
    def __init__(self, L):
        self.L = L

    def double(self, a):
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 0:
            a = a.reshape(1, 1)
        elif a.ndim == 1:
            a = a.reshape(1, -1)  # MATLAB row vector
        dims = (C.c_size_t * a.ndim)(*a.shape)
        f = np.asfortranarray(a)
        flat = np.ascontiguousarray(f.reshape(-1, order="F"))
        return self.L.fm_double(flat.ctypes.data_as(C.POINTER(C.c_double)), a.ndim, dims)

    def obj(self, cls, supers, **props):
        o = self.L.fm_object(cls.encode(), supers.encode())
        for k, v in props.items():
            self.L.fm_set_prop(o, k.encode(), self.double(v))
        return o

    def window(self, name, **props):
        return self.obj(f"windows.{name}", "windows.window", **props)

    def fnumber(self, cls, **props):
        return self.obj(cls, "f_numbers.f_number", **props)

    def normalization(self, name):
        return self.obj(f"normalizations.{name}", "normalizations.normalization")

    def single_complex(self, a):
        """complex single matrix (MATLAB layout) from a 2-D numpy array"""
        a = np.asarray(a)
        re = np.ascontiguousarray(a.real.reshape(-1, order="F"), dtype=np.float32)
        im = np.ascontiguousarray(a.imag.reshape(-1, order="F"), dtype=np.float32)
        fp = C.POINTER(C.c_float)
        return self.L.fm_single_complex(re.ctypes.data_as(fp), im.ctypes.data_as(fp), a.shape[0], a.shape[1])

    def empty(self):
        return self.L.fm_empty()

    def string(self, s):
        return self.L.fm_string(s.encode())

    def call(self, nlhs, args):
        prhs = (C.c_void_p * len(args))(*args)
        plhs = (C.c_void_p * max(nlhs, 1))()
        err = C.create_string_buffer(256)
        rc = self.L.fm_call(nlhs, plhs, len(args), prhs, err, 256)
        if rc:
            return None, err.value.decode()
        outs = []
        for i in range(nlhs):
            h = plhs[i]
            n = self.L.fm_numel(h)
            if self.L.fm_is_double(h):
                d = np.empty(n, dtype=np.float64)
                self.L.fm_get_double(h, d.ctypes.data_as(C.POINTER(C.c_double)))
                shape = tuple(self.L.fm_dim(h, j) for j in range(3))
                outs.append(d.reshape(shape, order="F"))
                self.L.fm_free(h)
                continue
            re = np.empty(n, dtype=np.float32)
            im = np.empty(n, dtype=np.float32)
            cplx = self.L.fm_is_single_complex(h)
            self.L.fm_get_single(h, re.ctypes.data_as(C.POINTER(C.c_float)),
                                 im.ctypes.data_as(C.POINTER(C.c_float)) if cplx else None)
            shape = (self.L.fm_dim(h, 0), self.L.fm_dim(h, 1))
            v = (re + 1j * im) if cplx else re
            outs.append(v.reshape(shape, order="F"))
            self.L.fm_free(h)
        return outs, None
