"""The C ABI: libgpe_hip.so loads, exports every entry point declared in
include/gpe_hip.h, and its host-only functions (time stepping, pump times)
agree with the oracle.  No device compute here."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import gpe_oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gpe_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpe_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree(ggp):
    from gpe_hip_amd import _lib

    decl = declared_functions()
    assert len(decl) >= 15
    assert set(decl) == set(_lib.SIGNATURES), set(decl) ^ set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol(ggp):
    lib = ggp.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.gpe_version() == 1


@pytest.mark.parametrize("dt,tspan,nsaves", [(0.01, (0, 0.4), 64), (0.05, (0, 3300), 512), (0.1, (0, 100), 256),
                                             (4, (0, 200), 1), (0.3 * 0.77, (0, 37.5), 123)])
def test_abi_time_stepping_matches_oracle(ggp, dt, tspan, nsaves):
    rdt, sps = ggp.resolve_fixed_timestepping(dt, tspan, nsaves)
    odt, osps, _ = orc.resolve_fixed_timestepping(dt, tspan, nsaves)
    assert rdt == odt and sps == osps
    times = ggp.pump_times(tspan, dt, nsaves)
    # oracle: the times at which the oracle evaluates the pump
    seen = []

    def pump(r, p, t):
        seen.append(t)
        return 0 * r[0]

    orc.solve(orc.Problem((np.zeros(16, complex),), (1.0,), pump=pump), tspan, dt, nsaves)
    assert np.array_equal(times, np.array(seen))


def test_abi_argument_errors(ggp):
    lib = ggp.load_library()
    out, sps = ctypes.c_double(), ctypes.c_int64()
    assert lib.gpe_resolve_timestepping(-1.0, 0.0, 1.0, 4, ctypes.byref(out), ctypes.byref(sps)) == -1
    assert b"dt" in lib.gpe_last_error()
    assert lib.gpe_resolve_timestepping(0.1, 0.0, 1.0, 0, ctypes.byref(out), ctypes.byref(sps)) == -1
    times = np.empty(5)
    assert lib.gpe_pump_times(0.0, 1.0, 0.1, 1, times.ctypes.data, 5) == -1  # needs 1 + 2*10 samples
    assert lib.gpe_create(None, None) == -1
    assert lib.gpe_destroy(None) == 0


def test_create_rejects_unsupported_shapes(ggp):
    """Shape checks run before any device call."""
    from gpe_hip_amd import _lib

    lib = ggp.load_library()
    # (ndim, n, ncomp, dtype): 4 dims; a length that is not a power of two,
    # too short or too long; 5 components; 3 components of 4096 points
    # (1536 threads per workgroup); 4 FP64 components of 2048 (LDS)
    for ndim, n, ncomp, dt in [(4, (16, 16, 16), 1, 1), (1, (100, 1, 1), 1, 1), (1, (8, 1, 1), 1, 1),
                               (1, (8192, 1, 1), 1, 1), (2, (64, 64, 1), 5, 1), (1, (4096, 1, 1), 3, 1),
                               (2, (64, 2048, 1), 4, 1)]:
        code = -2
        d = _lib.ProblemDesc()
        d.ndim, d.ncomp, d.batch, d.dtype = ndim, ncomp, 1, dt
        for i in range(3):
            d.n[i], d.L[i] = n[i], 1.0
        h = ctypes.c_void_p()
        assert lib.gpe_create(ctypes.byref(d), ctypes.byref(h)) == code
        assert not h.value


def test_create_without_device_reports_error(ggp):
    torch = pytest.importorskip("torch")
    if torch.cuda.device_count() > 0:
        pytest.skip("a HIP device is present")
    from gpe_hip_amd import _lib

    lib = ggp.load_library()
    d = _lib.ProblemDesc()
    d.ndim, d.ncomp, d.batch, d.dtype = 2, 1, 1, 1
    d.n[0] = d.n[1] = 64
    d.L[0] = d.L[1] = 1.0
    h = ctypes.c_void_p()
    assert lib.gpe_create(ctypes.byref(d), ctypes.byref(h)) == -3
    assert b"device" in lib.gpe_last_error()


def test_one_hip_runtime_when_the_library_loads_first():
    """Loading libgpe_hip.so before torch must not map a second HIP runtime
    (torch bundles its own libamdhip64); checked in a fresh interpreter."""
    import subprocess
    import sys

    code = (
        "import sys, re; sys.path.insert(0, %r)\n"
        "import __graft_entry__ as ge\n"
        "g = ge.load_package()\n"
        "from gpe_hip_amd import _lib\n"
        "_lib.load()\n"
        "maps = open('/proc/self/maps').read()\n"
        "print(len(set(re.findall(r'\\S*libamdhip64\\S*', maps))))\n" % ROOT
    )
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().splitlines()[-1] == "1"
