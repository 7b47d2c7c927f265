"""CPU-side checks of the C ABI: the library loads and exports every entry
point include/yahmc.h declares (no compute call is made without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "yahmc.h")
LIB = os.path.join(ROOT, "yahmc_amd", "lib", "libyahmc.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(yh_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("yh_create", "yh_Dphi", "yh_Hphi_sq", "yh_cg", "yh_bicgstab", "yh_cg_mshift", "yh_cg_mixed"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libyahmc.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libyahmc.so not built")
def test_python_mirror_binds_every_symbol():
    import yahmc_amd
    assert set(yahmc_amd.exported_symbols()) == set(declared())


@pytest.mark.skipif(not os.path.exists(LIB), reason="libyahmc.so not built")
def test_bad_lattice_rejected_before_any_device_call():
    import yahmc_amd
    with pytest.raises(yahmc_amd.YahmcError) as e:
        yahmc_amd.Lattice((3, 4, 4, 4))
    assert e.value.code == -1
    with pytest.raises(yahmc_amd.YahmcError) as e:
        yahmc_amd.Lattice((4, 4, 4, 4), nc=4, repr="adjoint")  # D = 15 is not built
    assert e.value.code == -6
