"""The C ABI library loads and exports every function include/ia.h declares;
host-side argument checking (no GPU compute needed)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared_functions():
    txt = (ROOT / "include" / "ia.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt):
        name = m.group(1)
        start = m.start()
        prev = txt[max(0, start - 80):start]
        if name in {"if", "for", "while", "sizeof", "return"}:
            continue
        # a declaration is preceded by a return type on the same statement
        if re.search(r"(void|ia_status|int|double|char\s*\*|const char \*)\s*\**\s*$", prev):
            names.add(name)
    return sorted(names)


def test_header_parses():
    names = _declared_functions()
    for n in ("esm", "esmgpu", "csvd", "tsvd", "to_pressure", "to_velocity", "sound_power",
              "sphere", "geometry", "cartesian_2_spherical", "ia_esm_sweep"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from inverseacoustics_amd import _lib
    L = _lib.lib()
    missing = [n for n in _declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the whole header
    assert set(_declared_functions()) <= set(_lib.SIGNATURES)


def test_version_and_air_roundtrip(ia):
    from inverseacoustics_amd import _lib
    assert b"gfx950" in _lib.lib().ia_version()
    ia.set_air(340.0, 1.2)
    assert ia.get_air() == (340.0, 1.2)
    with pytest.raises(ia.IAError):
        ia.set_air(-1.0, 1.2)
    ia.set_air(343.0, 1.21)


def test_host_only_utilities(ia, tmp_path):
    # cartesian_2_spherical (utils.cuh:414-427) and the MSH reader run on the host
    pts = np.array([[1.0, 0, 0], [0, -1.0, 0], [-1.0, 0, 1.0]])
    az = ia.cartesian_2_spherical(pts)
    np.testing.assert_allclose(az[0], [0.0, np.pi / 2, 1.0])
    np.testing.assert_allclose(az[1], [-np.pi / 2 + 2 * np.pi, np.pi / 2, 1.0])
    assert az[2][1] == pytest.approx(np.arctan2(1.0, 1.0) - 2 * np.pi)
    msh = tmp_path / "t.msh"
    msh.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 0 1 0\n"
                   "4 0 0 1\n$EndNodes\n$Elements\n4\n1 15 2 0 1 1\n2 2 2 0 1 1 2 3\n"
                   "3 2 2 0 1 1 3 3\n4 2 2 0 1 2 3 4\n$EndElements\n")
    r_ms, r_lh, e_n = ia.geometry(msh, 2 * np.pi, 0.5)
    assert r_ms.shape == (4, 3) and e_n.tolist() == [[0, 1, 2], [1, 2, 3]]
    np.testing.assert_allclose(r_lh, r_ms * 1.5)
    with pytest.raises(ia.IAError):
        ia.geometry(tmp_path / "missing.msh", 1.0, 0.0)


def test_sphere_host_formula_matches_oracle(ia):
    from inverseacoustics_amd import scenarios as sc
    from oracle import post
    r_s, r_h = sc.fibonacci_sphere(10, 0.1), sc.fibonacci_sphere(7, 0.4)
    got = ia.sphere(r_s, r_h, 12.0, 0.9, 0.1)
    ref = post.sphere(r_s, r_h, 12.0, 0.9, 0.1, 1.21, 343.0)
    for g, r in zip(got, ref):
        np.testing.assert_allclose(g, r, rtol=1e-13)


def test_compute_without_gpu_fails_loudly(ia):
    if ia.device_count() > 0:
        pytest.skip("GPU present")
    with pytest.raises(ia.IAError):
        ia.csvd(np.eye(3))


def test_quadrature_loader_reads_reference_tables():
    """ia_quadrature_load is host-only parsing: it reads the reference's text
    tables exactly as the oracle's loaders do."""
    from pathlib import Path

    import inverseacoustics_amd as ia
    from oracle import ibem as ib

    gold = Path(__file__).parent / "golden"
    rule, pages = ia.quadrature_load(gold / "IPF.txt", gold / "IPFs.txt")
    assert np.array_equal(rule, ib.load_ipf(gold / "IPF.txt"))
    assert np.array_equal(pages, ib.load_ipfs(gold / "IPFs.txt"))
    with pytest.raises(ia.IAError):
        ia.quadrature_load(gold / "missing.txt", gold / "IPFs.txt")
