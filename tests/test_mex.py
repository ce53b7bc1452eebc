"""The MEX gateway (sid_amd/mex/sid_mex_gateway.c) compiled against a test
stand-in of MATLAB's mex.h and called without a GPU: usage and argument
errors map to 'sid:arg', the device calls to 'sid:nodev', before any array is
touched on a device."""
import numpy as np
import pytest

from tests._mex import Mex, MexError


@pytest.fixture(scope="module", autouse=True)
def built():
    from tests.mex_stub.build_mex import build
    build()


def _err(fn, *args, nlhs=1):
    with pytest.raises(MexError) as e:
        fn(*args, nlhs=nlhs)
    return e.value


H = np.ones((7, 7, 3, 3, 2), dtype=np.float32)


@pytest.mark.parametrize("name,args", [
    ("forwardProjectGPU", (H,)),
    ("backwardProjectGPU_new", (H,)),
    ("forwardProjectACC", (H, np.ones((9, 9)))),
    ("backwardProjectACC_new", (H, np.ones((9, 9)))),
    ("sidDeconvRL", (H, np.ones((9, 9, 2)))),
    ("sidLSDeconv", (H,)),
    ("generate_LFM_library_GPU", ([np.ones((9, 9, 2))], np.ones((1, 3)))),
    ("generate_LFM_library_CPU", (np.ones((1, 3)),)),
    ("sidPSF", ()),
])
def test_usage_errors(name, args):
    e = _err(Mex(name), *args)
    assert e.ident == "sid:arg" and "usage" in e.msg


def test_argument_errors_before_any_device_call():
    e = _err(Mex("sidDeconvRL"), H, np.ones((9, 9, 2), np.float32), 3.0, np.ones((9, 8, 2), np.float32))
    assert e.ident == "sid:arg" and "Xguess" in e.msg
    e = _err(Mex("sidDeconvRL"), H, np.ones((9, 9, 2), np.float32), 2.5, np.ones((9, 9, 2), np.float32))
    assert e.ident == "sid:arg" and "maxIter" in e.msg
    e = _err(Mex("sidLSDeconv"), H, np.ones((9, 9, 2)), np.ones(3))
    assert e.ident == "sid:arg" and "struct" in e.msg
    e = _err(Mex("generate_LFM_library_GPU"), np.ones((9, 9, 2)), np.ones((1, 3)), np.ones((1, 1)), H,
             {"image_size": np.array([9.0, 9.0]), "axial": 1.0, "neur_rad": 2.0})
    assert e.ident == "sid:arg" and "cell" in e.msg
    e = _err(Mex("generate_LFM_library_CPU"), np.ones((2, 2)), {"H": H}, 2.0, np.ones(3), np.array([9.0, 9.0, 2.0]))
    assert e.ident == "sid:arg" and "N x 3" in e.msg
    e = _err(Mex("sidPSF"), "destroy", np.array([12345], dtype=np.uint64))
    assert e.ident == "sid:arg" and "live" in e.msg
    e = _err(Mex("sidPSF"), "resize")
    assert e.ident == "sid:arg"
    e = _err(Mex("sidPSF"), "kernel", np.array([12345], dtype=np.uint64), np.ones((3, 3, 1)))
    assert e.ident == "sid:arg" and "live" in e.msg
    e = _err(Mex("sidPSF"), "kernel", np.array([12345], dtype=np.uint64))
    assert e.ident == "sid:arg" and "usage" in e.msg
    e = _err(Mex("forwardProjectGPU"), np.array([7], dtype=np.uint64), np.ones((9, 9)))
    assert e.ident == "sid:arg" and "live" in e.msg
    e = _err(Mex("forwardProjectACC"), H, np.ones((9, 9)), np.ones((3, 2)))
    assert e.ident == "sid:arg" and "CAindex" in e.msg


@pytest.mark.parametrize("name,args", [
    ("forwardProjectGPU", (H, np.ones((9, 9, 2)))),
    ("backwardProjectGPU_new", (H, np.ones((9, 9)))),
    ("sidDeconvRL", (H, np.ones((9, 9, 2)), 2.0, np.ones((9, 9, 2)))),
    ("sidPSF", ("create", H)),
    # ker_shape alone selects no kernel (generate_LFM_library_GPU.m:50-58: only ker_param does)
    ("generate_LFM_library_GPU", ([np.ones((9, 9, 2))], np.ones((1, 3)), np.ones((1, 1)), H,
                                  {"image_size": np.array([9.0, 9.0]), "axial": 1.0, "neur_rad": 2.0,
                                   "ker_shape": "ball"})),
    ("generate_LFM_library_GPU", ([np.ones((9, 9, 2))], np.ones((1, 3)), np.ones((1, 1)), H,
                                  {"image_size": np.array([9.0, 9.0]), "axial": 1.0, "neur_rad": 2.0,
                                   "ker_shape": "ball", "ker_param": np.array([2.0, 1.0])})),
])
def test_device_calls_report_nodev_without_a_gpu(name, args):
    import sid_amd
    if sid_amd.device_count() > 0:
        pytest.skip("a gfx950 device is visible")
    e = _err(Mex(name), *args)
    assert e.ident == "sid:nodev"
