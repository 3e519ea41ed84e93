"""C-ABI and host-mirror checks that need no GPU.

- libhoe.so loads and exports every entry point include/hoe.h declares;
- hoe_create validates its arguments before touching a device and maps
  failures to the reference's exception types (ArgumentError for bad
  keywords / term counts / recon_terms, DimensionMismatch for sizes:
  explicit_gpu_operator.jl:299-330, prep_kspha.jl:12-21, prep_kspha.jl:87-92);
- the normal-operator weight rules (explicit_gpu_operator.jl:405-415);
- the product never falls back to the CPU: a missing library raises.
"""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import hoe
from hoe import _lib
from hoe.operator import _check_diag

ROOT = Path(__file__).resolve().parent.parent


def _header_symbols():
    text = (ROOT / "include" / "hoe.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hoe_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = hoe.load()
    syms = _header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.EXPORTS)
    assert lib.hoe_version().decode().startswith("libhoe")


def test_device_count_never_fails():
    assert hoe.load().hoe_device_count() >= 0


def _desc(n_term=16, n_sam=5, recon_terms=None, n_cha=1, csm=False):
    keep = []
    n = 4 * 3 * 2
    x = np.linspace(-1, 1, n).astype(np.float32)
    k = np.zeros((n_sam, n_term), np.float32)
    t = np.zeros(n_sam, np.float32)
    keep += [x, k, t]
    d = _lib.ModelDesc()
    d.n_x, d.n_y, d.n_z = 4, 3, 2
    d.x = d.y = d.z = x.ctypes.data
    d.n_term, d.n_sam, d.n_cha = n_term, n_sam, n_cha
    d.kspha, d.times = k.ctypes.data, t.ctypes.data
    if csm:
        c = np.ones((n_cha, n), np.complex64)
        keep.append(c)
        d.csm = c.ctypes.data
    if recon_terms is not None:
        d.recon_terms = recon_terms.encode()
    return d, keep


def _create(d):
    lib = hoe.load()
    h = C.c_void_p()
    rc = lib.hoe_create(C.byref(d), C.byref(h))
    msg = lib.hoe_last_error(None).decode()
    if rc == 0:
        lib.hoe_destroy(h)
    return rc, msg


def test_create_validates_before_device_use():
    lib = hoe.load()
    assert lib.hoe_create(None, C.byref(C.c_void_p())) == _lib.HOE_EINVAL
    d, keep = _desc(n_term=12)
    assert _create(d) == (_lib.HOE_EINVAL, "kspha requires 9 or 16 terms")
    d, keep = _desc(recon_terms="0121")
    rc, msg = _create(d)
    assert rc == _lib.HOE_EINVAL and "only '0' and '1'" in msg
    d, keep = _desc(recon_terms="111")
    rc, msg = _create(d)
    assert rc == _lib.HOE_EINVAL and "4 binary digits" in msg
    d, keep = _desc(n_term=9, recon_terms="1111")
    rc, msg = _create(d)
    assert rc == _lib.HOE_EINVAL and "3 binary digits" in msg
    d, keep = _desc(n_cha=3, csm=False)
    assert _create(d)[0] == _lib.HOE_EDIM
    d, keep = _desc()
    d.backend = 7
    assert _create(d)[0] == _lib.HOE_EINVAL
    d, keep = _desc(n_sam=0)
    assert _create(d)[0] == _lib.HOE_EINVAL


@pytest.mark.skipif(hoe.load().hoe_device_count() > 0, reason="checks the no-device path")
def test_create_without_device_reports_enodev():
    d, keep = _desc()
    rc, msg = _create(d)
    assert rc == _lib.HOE_ENODEV and "no HIP device" in msg
    with pytest.raises(hoe.NoDeviceError):
        hoe.HighOrderExplicitOp(hoe.Grid(4, 3, 2, 0.1, 0.1, 0.1), np.zeros((16, 5), np.float32),
                                np.zeros(5, np.float32))


def test_host_mirror_argument_errors():
    g = hoe.Grid(4, 3, 2, 0.1, 0.1, 0.1)
    k = np.zeros((16, 5), np.float32)
    t = np.zeros(5, np.float32)
    with pytest.raises(hoe.ArgumentError):
        hoe.HighOrderExplicitOp(g, k, t, implementation="array")
    with pytest.raises(hoe.ArgumentError):
        hoe.HighOrderExplicitOp(g, k, t, implementation="other")
    with pytest.raises(AssertionError):                 # array_operator.jl:160
        hoe.HighOrderExplicitOp(g, k, t, kspha_dt=k[:, :-1])
    with pytest.raises(hoe.ArgumentError):              # test_explicit_interface.jl:71
        hoe.HighOrderExplicitOp(g, k, t, backend="blocked", kspha_dt=k)
    with pytest.raises(hoe.ArgumentError):
        hoe.HighOrderExplicitOp(g, k, t, backend="invalid")
    with pytest.raises(hoe.ArgumentError):
        hoe.HighOrderExplicitOp(g, k, t, gpus=[])
    with pytest.raises(hoe.ArgumentError):
        hoe.HighOrderExplicitOp(g, k, t, gpus=[0, 0])
    with pytest.raises(hoe.ArgumentError):
        hoe.HighOrderExplicitOp(g, np.zeros((12, 5), np.float32), t)
    with pytest.raises(hoe.DimensionMismatch):
        hoe.HighOrderExplicitOp(g, k, np.zeros(4, np.float32))
    with pytest.raises(hoe.DimensionMismatch):
        hoe.HighOrderExplicitOp(g, k, t, mask=np.ones(5, bool))
    with pytest.raises(hoe.DimensionMismatch):
        hoe.HighOrderExplicitOp(g, k, t, fieldmap=np.zeros(7, np.float32))
    with pytest.raises(hoe.DimensionMismatch):
        hoe.HighOrderExplicitOp(g, k, t, csm=np.ones((23, 2), np.complex64))
    with pytest.raises(hoe.DimensionMismatch):
        hoe.HighOrderExplicitOp(g, k, t, concomitant=np.zeros((3, 5), np.float32))
    with pytest.raises(hoe.ArgumentError):
        hoe.HighOrderExplicitOp(g, k, t, concomitant=np.zeros((4, 5), np.complex64))


def test_normal_weight_rules():
    ns, nrow = 4, 8
    np.testing.assert_array_equal(_check_diag(np.ones(ns), nrow, ns), np.ones(ns, np.float32))
    _check_diag(np.ones(nrow), nrow, ns)
    _check_diag(np.ones(ns) + 1e-9j, nrow, ns)      # imaginary part within 10 eps
    with pytest.raises(hoe.ArgumentError):
        _check_diag(np.full(ns, 1 + 1j), nrow, ns)
    with pytest.raises(hoe.ArgumentError):
        _check_diag(-np.ones(ns), nrow, ns)
    with pytest.raises(hoe.ArgumentError):
        _check_diag(np.array([1.0, np.nan, 1.0, 1.0]), nrow, ns)
    with pytest.raises(hoe.DimensionMismatch):
        _check_diag(np.ones(5), nrow, ns)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(hoe.LibraryNotBuilt):
        _lib.load(tmp_path / "libhoe.so")


def test_product_does_not_reference_the_oracle():
    pkg = ROOT / "highordermri.jl_amd"
    for f in list(pkg.rglob("*.py")) + list(pkg.rglob("*.hip")) + list(pkg.rglob("*.cpp")) + \
            list(pkg.rglob("*.h")) + list(pkg.rglob("Makefile")):
        text = f.read_text()
        assert "import oracle" not in text and "liboracle" not in text and "hoe_oracle" not in text, f


def test_out_vector_checks_on_host_arrays():
    """The output-vector checks of the operator wrapper (host arrays, no GPU needed)."""
    import numpy as np
    from hoe import operator as opm
    v = opm._Vec(np.zeros(6, np.complex64), 6, "x")
    assert opm._out(None, v, 4).shape == (4,)
    ok = np.empty(4, np.complex64)
    assert opm._out(ok, v, 4) is ok
    with pytest.raises(hoe.DimensionMismatch):
        opm._out(np.empty(5, np.complex64), v, 4)
    with pytest.raises(hoe.ArgumentError):
        opm._out(np.empty(4, np.complex128), v, 4)
    with pytest.raises(hoe.ArgumentError):
        opm._out(np.empty(8, np.complex64)[::2], v, 4)
    ro = np.empty(4, np.complex64)
    ro.flags.writeable = False
    with pytest.raises(hoe.ArgumentError):
        opm._out(ro, v, 4)
    with pytest.raises(hoe.ArgumentError):
        opm._out([0j] * 4, v, 4)
