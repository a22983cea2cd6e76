"""Text writers and printers with the format of CUDA/utils.cuh:32-197 (host
code, no GPU).  Expected outputs are literal strings worked out from the C
conversions the reference uses ("%f" rounds to six decimals, "%e" prints
d.dddddde+XX)."""
import ctypes as C
import subprocess
import sys

import numpy as np
import pytest


def test_write_complex_and_append(ia, tmp_path):
    f = tmp_path / "p.txt"
    ia.write_file_complex("w", f, [1.5 - 2.25j, 1e-7 + 3j, -0.0000005 + 123456.7890125j])
    assert f.read_text() == "1.500000,-2.250000;0.000000,3.000000;-0.000000,123456.789012\n"
    ia.write_file_complex("a", f, [2j])
    assert f.read_text().splitlines()[1] == "0.000000,2.000000"
    ia.write_file_complex("w", f, [])
    assert f.read_text() == "\n"


def test_write_double_and_string(ia, tmp_path):
    f = tmp_path / "w.txt"
    ia.write_file_double("w", f, [0.1, -2.0, 1e20, 0.0000004])
    assert f.read_text() == "0.100000;-2.000000;100000000000000000000.000000;0.000000\n"
    ia.write_file_string("a", f, "k=18.3\n")
    assert f.read_text().endswith("0.000000\nk=18.3\n")


def test_write_matrix_row_major_text(ia, tmp_path):
    f = tmp_path / "m.txt"
    A = np.array([[1 + 2j, 3 - 4j, 5.0], [-1j, 0.5, 2 + 0.25j]])   # 2 x 3
    ia.write_matrix("f", f, A)
    assert f.read_text() == ("1.000000,2.000000,3.000000,-4.000000,5.000000,0.000000,\n"
                             "-0.000000,-1.000000,0.500000,0.000000,2.000000,0.250000,\n\n")
    ia.write_matrix("e", f, A[:1, :1])
    assert f.read_text() == "1.000000e+00,2.000000e+00,\n\n"


def test_legacy_writer_names(tmp_path):
    from inverseacoustics_amd import _lib
    L = _lib.lib()
    f = str(tmp_path / "legacy.txt").encode()
    v = np.array([1 + 1j, 2 - 0.5j])
    L.write_file_cuDoubleComplex(b"w", f, _lib.zptr(v), 2)
    d = np.array([4.0])
    L.write_file_double(b"a", f, _lib.dptr(d), 1)
    L.write_file_string(b"a", f, b"end")
    assert (tmp_path / "legacy.txt").read_text() == "1.000000,1.000000;2.000000,-0.500000\n4.000000\nend"


def test_unwritable_file_is_an_io_error(ia, tmp_path):
    with pytest.raises(ia.IAError) as e:
        ia.write_file_double("w", tmp_path / "no" / "such" / "dir.txt", [1.0])
    assert e.value.status == 6
    with pytest.raises(ia.IAError):
        ia.write_file_double("r", tmp_path / "x.txt", [1.0])   # not an output mode


def test_printers_stdout():
    # run in a child so the C stdio buffer is captured
    code = (
        "import numpy as np\n"
        "from inverseacoustics_amd import _lib\n"
        "L=_lib.lib()\n"
        "a=np.asfortranarray(np.array([[1+2j,3j],[0.5,-1]]))\n"
        "L.print_matrix_cuDoubleComplex(b'f',_lib.zptr(a),2,2)\n"
        "d=np.asfortranarray(np.array([[1.0,2.5],[3.25,-4.0]]))\n"
        "L.print_matrix_double(b'f',_lib.dptr(d),2,2)\n"
        "L.print_matrix_double(b'e',_lib.dptr(d),1,2)\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                         check=True).stdout
    z = "%.16f %.16fi  "
    expect = ((z % (1, 2)) + (z % (0, 3)) + "\n" + (z % (0.5, 0)) + (z % (-1, 0)) + "\n\n"
              "1.0000 2.5000 \n3.2500 -4.0000 \n\n"
              "1e+00f 3e+00f \n\n")   # r = 1 reads a[0], a[1] (leading dimension r)
    assert out == expect
