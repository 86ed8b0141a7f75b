"""The C-ABI library builds, loads and exports every entry point include/mdforce.h declares
(no compute calls: no GPU needed)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "mdforce.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mdf_[a-z_]+)\s*\(", txt)))


def test_library_exports_declared_symbols():
    from md_sim_amd import engine
    L = engine.lib()
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n


def test_last_error_callable_without_device():
    from md_sim_amd import engine
    L = engine.lib()
    assert isinstance(L.mdf_last_error(), bytes)


def test_create_without_device_fails_loudly():
    """On a machine without a HIP device the engine refuses to run (no CPU fallback)."""
    import pytest
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from md_sim_amd import systems
    from md_sim_amd.engine import ForceEngine, MDForceError
    with pytest.raises(MDForceError) as e:
        ForceEngine(systems.config("pore_small"))
    assert e.value.code in (2, 7)
