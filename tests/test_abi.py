"""The C-ABI library: it builds, loads, exports every declared symbol and
validates arguments before touching the device (runs without a GPU)."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from finmath_hip import _abi

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
HEADERS = [os.path.join(INCLUDE, h) for h in sorted(os.listdir(INCLUDE)) if h.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(fm_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declarations_match_binding_list():
    assert declared_functions() == sorted(_abi.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    path = _abi.library_path()
    assert os.path.exists(path), "library not built"
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared_functions() if s not in exported]
    assert not missing, missing


def test_library_loads_and_binds():
    lib = _abi.load()
    for s in _abi.EXPORTED_SYMBOLS:
        assert hasattr(lib, s)


def _ctx(**kw):
    args = dict(times=np.arange(5) * 0.5, tenor=np.arange(5) * 0.5, num_factors=2, num_paths=16, seed=3141)
    args.update(kw)
    return _abi.LmmContext(**args)


@pytest.mark.parametrize("kw,exc", [
    (dict(num_factors=0), _abi.IllegalArgumentException),
    (dict(num_paths=0), _abi.IllegalArgumentException),
    (dict(times=np.array([0.0, 0.5, 0.5])), _abi.IllegalArgumentException),
    (dict(tenor=np.array([0.0, 1.0, 0.5])), _abi.IllegalArgumentException),
    (dict(num_factors=11), _abi.UnsupportedOperationException),
    (dict(tenor=np.arange(82) * 0.5), _abi.UnsupportedOperationException),
    (dict(libor_cap=math.nan), _abi.IllegalArgumentException),
    (dict(times=np.array([0.25, 0.5, 1.0])), _abi.UnsupportedOperationException),
])
def test_invalid_configuration_raises_before_device_use(kw, exc):
    with pytest.raises(exc):
        _ctx(**kw)


def test_reference_validation_message():
    # BrownianMotionFromMersenneRandomNumbers.java:77-78 wording
    with pytest.raises(_abi.IllegalArgumentException, match="Number of factors must be greater or equal 1"):
        _ctx(num_factors=0)


@pytest.mark.parametrize("blob", [b"", b"garbage-bytes-not-a-context", b"FMLMMSER\x63\x00\x00\x00" + bytes(64),
                                  b"FMLMMSER\x01\x00\x00\x00" + bytes(8)])
def test_malformed_serialised_context_is_rejected_before_device_use(blob):
    if not blob:
        with pytest.raises(_abi.IllegalArgumentException):
            _abi.check(_abi.load().fm_lmm_deserialize(None, 0, None))
        return
    with pytest.raises(_abi.IllegalArgumentException):
        _abi.LmmContext.deserialize(blob)
