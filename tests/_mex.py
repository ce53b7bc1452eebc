"""ctypes harness around the MEX gateway builds of tests/mex_stub (test
infrastructure): builds MATLAB-like arrays, calls mexFunction, reads outputs."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "mex_stub", "_build")
CLS = {np.dtype(np.float64): 6, np.dtype(np.float32): 7, np.dtype(np.int32): 12, np.dtype(np.uint64): 15}
CLS_NAME = {6: "double", 7: "single", 12: "int32", 15: "uint64", 4: "char", 2: "struct", 1: "cell"}
_libs = {}


def _load(name):
    if name not in _libs:
        path = os.path.join(BUILD, name + ".so")
        if not os.path.exists(path):
            from tests.mex_stub.build_mex import build
            build()
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.stub_numeric.restype = vp
        L.stub_numeric.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), vp]
        L.stub_char.restype = vp
        L.stub_char.argtypes = [ctypes.c_char_p]
        L.stub_struct.restype = vp
        L.stub_struct.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(vp)]
        L.stub_cell.restype = vp
        L.stub_cell.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        for f in ("stub_class", "stub_is_sparse", "stub_ndim"):
            getattr(L, f).restype = ctypes.c_int
            getattr(L, f).argtypes = [vp]
        L.stub_dim.restype = ctypes.c_int64
        L.stub_dim.argtypes = [vp, ctypes.c_int]
        for f in ("stub_data", "stub_ir", "stub_jc"):
            getattr(L, f).restype = vp
            getattr(L, f).argtypes = [vp]
        L.stub_err_id.restype = ctypes.c_char_p
        L.stub_err_msg.restype = ctypes.c_char_p
        L.stub_call.restype = ctypes.c_int
        L.stub_call.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
        L.stub_set_matlab.restype = None
        L.stub_set_matlab.argtypes = [_MATLAB_CB]
        _libs[name] = L
        L.stub_set_matlab(_make_cb(L))
    return _libs[name]


# MATLAB functions the gateways call back through mexCallMATLAB (name -> python
# callable taking and returning what tests/_mex.Mex converts).
MATLAB_FUNCTIONS = {}
_MATLAB_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                              ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p)
_callbacks = []   # keeps the ctypes callbacks alive


def _make_cb(L):
    def cb(nlhs, plhs, nrhs, prhs, name):
        fn = MATLAB_FUNCTIONS.get(name.decode())
        if fn is None:
            return 1
        m = Mex.__new__(Mex)
        m.L = L
        args = [m.out(prhs[i])[1] for i in range(nrhs)]
        res = fn(*args)
        if nlhs >= 1:
            plhs[0] = m.arr(res)
        return 0
    c = _MATLAB_CB(cb)
    _callbacks.append(c)
    return c


class MexError(Exception):
    def __init__(self, ident, msg):
        super().__init__(f"{ident}: {msg}")
        self.ident = ident
        self.msg = msg


class Mex:
    """One MEX entry point (e.g. Mex('forwardProjectGPU')(H, V))."""

    def __init__(self, name):
        self.L = _load(name)

    def arr(self, x):
        L = self.L
        if isinstance(x, str):
            return L.stub_char(x.encode())
        if isinstance(x, dict):
            names = (ctypes.c_char_p * len(x))(*[k.encode() for k in x])
            vals = (ctypes.c_void_p * len(x))(*[self.arr(v) for v in x.values()])
            return L.stub_struct(len(x), names, vals)
        if isinstance(x, list):
            vals = (ctypes.c_void_p * len(x))(*[self.arr(v) for v in x])
            return L.stub_cell(len(x), vals)
        a = np.asarray(x)
        if a.dtype not in CLS:
            a = a.astype(np.float64)
        if a.ndim == 0:
            a = a.reshape(1, 1)
        elif a.ndim == 1:
            a = a.reshape(1, -1)
        a = np.asfortranarray(a)
        dims = (ctypes.c_int64 * a.ndim)(*a.shape)
        return L.stub_numeric(CLS[a.dtype], a.ndim, dims, a.ctypes.data_as(ctypes.c_void_p))

    def out(self, p):
        L = self.L
        cls = L.stub_class(p)
        shape = tuple(L.stub_dim(p, i) for i in range(L.stub_ndim(p)))
        if L.stub_is_sparse(p):
            import scipy.sparse
            m, n = shape
            jc = np.ctypeslib.as_array(ctypes.cast(L.stub_jc(p), ctypes.POINTER(ctypes.c_size_t)), (n + 1,)).copy()
            nnz = int(jc[-1])
            ir = np.ctypeslib.as_array(ctypes.cast(L.stub_ir(p), ctypes.POINTER(ctypes.c_size_t)), (max(nnz, 1),))[:nnz]
            pr = np.ctypeslib.as_array(ctypes.cast(L.stub_data(p), ctypes.POINTER(ctypes.c_double)), (max(nnz, 1),))[:nnz]
            return "sparse", scipy.sparse.csc_matrix((pr.copy(), ir.copy(), jc), shape=(m, n))
        if cls == 4:   # char: UTF-16 code units
            n = int(np.prod(shape))
            buf = np.ctypeslib.as_array(ctypes.cast(L.stub_data(p), ctypes.POINTER(ctypes.c_uint16)), (max(n, 1),))[:n]
            return "char", "".join(chr(c) for c in buf)
        dt = {6: np.float64, 7: np.float32, 12: np.int32, 15: np.uint64}[cls]
        n = int(np.prod(shape))
        buf = np.ctypeslib.as_array(ctypes.cast(L.stub_data(p), ctypes.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                    (max(n, 1),))[:n]
        return CLS_NAME[cls], buf.copy().reshape(shape, order="F")

    def __call__(self, *args, nlhs=1):
        prhs = (ctypes.c_void_p * max(len(args), 1))(*[self.arr(a) for a in args])
        plhs = (ctypes.c_void_p * max(nlhs, 1))()
        if self.L.stub_call(nlhs, plhs, len(args), prhs):
            raise MexError(self.L.stub_err_id().decode(), self.L.stub_err_msg().decode())
        outs = [self.out(plhs[i]) for i in range(nlhs) if plhs[i]]
        return outs[0] if nlhs == 1 else outs

    def exit(self):
        self.L.stub_exit()
