"""Gauge configuration files: the library's writer and reader against an
independent restatement of the two formats with Python's struct module
(native: include/gaugefield.h:20-36 header_t + library/gaugefield.cpp:44-179
+ suNg::write, suNg.cpp:403-412; HiRep: library/hirep.cpp:56-212, big
endian).  Host only: no device is involved."""
import os
import struct

import numpy as np
import pytest

yh = pytest.importorskip("yahmc_amd")
if not os.path.exists(yh.LIBPATH):
    pytest.skip("libyahmc.so not built", allow_module_level=True)

from fields import eo_geometry, random_gauge  # noqa: E402


def lex_to_eo(dims):
    T, X, Y, Z = dims
    V = T * X * Y * Z
    out = []
    for t in range(T):
        for x in range(X):
            for y in range(Y):
                for z in range(Z):
                    lex = ((t * X + x) * Y + y) * Z + z
                    out.append(((t + x + y + z) & 1) * (V // 2) + lex // 2)
    return out


def expected_body(links, dims, nc, endian):
    parts = []
    n2 = nc * nc
    for id_ in lex_to_eo(dims):
        for mu in range(4):
            u = links[4 * id_ + mu]
            for e in range(n2):
                parts.append(struct.pack(endian + "dd", u[e], u[n2 + e]))
    return b"".join(parts)


@pytest.mark.parametrize("nc", [2, 3])
def test_native_format(tmp_path, nc):
    dims = (4, 2, 2, 4)
    rng = np.random.default_rng(nc)
    links = random_gauge(rng, int(np.prod(dims)), nc)
    path = tmp_path / "cfg.bin"
    yh.config_write(path, links, dims, nc, repr=2, plaquette=0.625, num_cnfg=17, num_accept=11)
    raw = path.read_bytes()
    assert len(raw) == 256 + 4 * int(np.prod(dims)) * nc * nc * 16
    head = struct.unpack("=10id", raw[:48])
    assert head[:10] == (0x47464359, 0x20, *dims, nc, 2, 17, 11)
    assert head[10] == 0.625
    assert raw[48:256] == bytes(208)
    assert raw[256:] == expected_body(links, dims, nc, "=")
    info, back = yh.config_read(path)
    assert info["dims"] == dims and info["nc"] == nc and info["repr"] == 2
    assert (info["num_cnfg"], info["num_accept"], info["plaquette"]) == (17, 11, 0.625)
    assert np.array_equal(back, links)


def test_hirep_format(tmp_path):
    dims = (2, 2, 4, 4)
    nc = 3
    rng = np.random.default_rng(5)
    links = random_gauge(rng, int(np.prod(dims)), nc)
    path = tmp_path / "run1n42"
    yh.config_write(path, links, dims, nc, plaquette=0.5, format="hirep")
    raw = path.read_bytes()
    assert struct.unpack(">5id", raw[:28]) == (nc, *dims, 0.5)
    assert raw[28:] == expected_body(links, dims, nc, ">")
    info, back = yh.config_read(path, format="hirep")
    assert info["num_cnfg"] == 42 and info["repr"] == -1 and info["plaquette"] == 0.5
    assert np.array_equal(back, links)


def test_bad_files_rejected(tmp_path):
    path = tmp_path / "bad.bin"
    path.write_bytes(bytes(256))
    with pytest.raises(yh.YahmcError):
        yh.config_read(path)
    dims = (2, 2, 2, 2)
    links = random_gauge(np.random.default_rng(1), 16, 3)
    yh.config_write(path, links, dims, 3)
    path.write_bytes(path.read_bytes()[:-8])  # truncated
    with pytest.raises(yh.YahmcError):
        yh.config_read(path)
