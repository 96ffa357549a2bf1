"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/khrr.h declares, the Python binding covers them all, and host-only
entry points (constants, argument validation) behave.  No GPU compute."""
import ctypes
import math
import subprocess

import pytest

import khrr


def test_library_exports_every_declared_symbol():
    names = khrr.header_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", khrr.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_binding_covers_header():
    assert set(khrr.header_functions()) == set(khrr._SIGS)
    lib = khrr.lib()
    for name in khrr.header_functions():
        assert getattr(lib, name) is not None


def test_abi_version():
    assert khrr.lib().khrr_abi_version() == 2  # the superblock field of khrr_tune


def test_extpot_constants_match_reference_formulas():
    # nbody.cu:1111-1120
    Mh, a, Rh, Mb, b, Rb = 2.0, 0.3333, 1.0, 0.5, 0.0431, 0.2
    e = khrr.extpot_init(Mh, a, Rh, Mb, b, Rb)
    Rh2 = 3.0 * Rh
    rcore1 = Rh / a
    root1 = math.sqrt(1.0 + Rb * Rb / (b * b))
    assert e.Rh2 == Rh2
    assert e.con == pytest.approx(Mh / (rcore1 - math.atan(rcore1)), rel=1e-15)
    assert e.const1 == pytest.approx(Mb / (b * math.log(Rb / b + root1) - Rb / root1), rel=1e-14)
    assert e.Mh_inf == pytest.approx(Mh * (Rh2 / a - math.atan(Rh2 / a)) /
                                     (rcore1 - math.atan(rcore1)), rel=1e-15)


def test_extpot_force_host_matches_oracle():
    import oracle as orc
    e = khrr.extpot_init(2.0, 0.3333, 1.0, 0.5, 0.0431, 0.2)
    o = orc.extpot(2.0, 0.3333, 1.0, 0.5, 0.0431, 0.2)
    for p in ([0.1, 0.0, 0.05], [0.0, 0.0, 0.0], [3.5, -1.0, 0.2], [0.15, 0.1, -0.02]):
        assert khrr.extpot_force(e, p) == pytest.approx(list(orc.dev_fex(o, p)), rel=1e-15,
                                                        abs=1e-300)


def test_context_without_device_reports_error():
    # In the build container there is no GPU: creation must fail with a
    # message, not crash or exit.
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = khrr.lib().khrr_ctx_create(0, ctypes.byref(h))
    assert rc == -2 and not h.value
    assert b"device" in khrr.lib().khrr_last_error(None)


def test_null_context_is_rejected():
    lib = khrr.lib()
    assert lib.khrr_ctx_sync(None) == -1
    assert b"null context" in lib.khrr_last_error(None)
    assert lib.khrr_accel(None, 1, None, None, 1, None, None, None, None, None, 0, None) == -1


def test_python_binding_raises_when_library_missing(tmp_path):
    with pytest.raises(ImportError):
        khrr.load_library(str(tmp_path / "libkhrr.so"))
