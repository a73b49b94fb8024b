"""The reference's input formats through the library (include/tpls_io.h);
host code only, so these run without a GPU.

Pins:
* the known answers of the reference's own tests: test/io_test.f90:71-165
  (data (i+1)(j+2)(k+3) on bounds 1:9, 0:10, 1:11; mismatch codes 1, 2, 2, 3,
  3) and test/options_utils_test.f90:49-434 (every test ported);
* the reference's options_utils.F90 itself: the fixtures of
  tests/golden/make_options_golden.py (its load_options and typed getters on
  an edge-case file, its put_*_option and save_options output), and the live
  module where oracle/_ref is built;
* the Fortran runtime for the data format: a file written by Fortran
  list-directed output (tests/fortran/drive_io.f90, the statements of
  save_data) reads back bit-exactly, and a file the library writes reads
  bit-exactly with Fortran's list-directed reads (those of load_data);
* the spellings other compilers' list-directed output uses (gfortran's
  17-digit fields, D exponents, letterless exponents, repeat counts, null
  values), checked against correctly rounded values and against how the
  Fortran runtime reads the same records.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from tpls_amd import _lib, io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
DRIVE_IO = os.path.join(ROOT, "tests", "fortran", "drive_io")
REF_OPTIONS = os.path.join(ROOT, "oracle", "_ref", "ref_options_driver")

pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")


def bits(a):
    return np.asarray(a, dtype=np.float64).view(np.int64)


# ---------------------------------------------------------------------------
# io_test.f90

LX, LY, MX, MY, NX, NY = 1, 9, 0, 10, 1, 11


def out_data():
    """populate_out_data of io_test.f90:55-67."""
    i, j, k = np.meshgrid(np.arange(LX, LY + 1), np.arange(MX, MY + 1), np.arange(NX, NY + 1), indexing="ij")
    return np.asfortranarray(((i + 1) * (j + 2) * (k + 3)).astype(np.float64))


def test_save_data(tmp_path):
    f = tmp_path / "io_test.tmp"
    io.save_data(f, "Test data", out_data(), LX, LY, MX, MY, NX, NY)
    assert f.exists()
    lines = f.read_text().splitlines()
    assert lines[0].split() == ["DATA:", "Test", "data"]
    assert lines[1].split() == ["DIMENSIONS:", "Lmin", "Lmax", "Mmin", "Mmax", "Nmin", "Nmax"]
    assert [int(x) for x in lines[2].split()] == [LX, LY, MX, MY, NX, NY]
    # one record per value, i fastest (io.F90:41-46)
    vals = [float(x) for x in lines[3:]]
    assert vals == list(out_data().ravel(order="F"))


def test_load_data(tmp_path):
    f = tmp_path / "io_test.tmp"
    d = out_data()
    io.save_data(f, "Test data", d, LX, LY, MX, MY, NX, NY)
    data, lx, ly, mx, my, nx, ny, ierr = io.load_data(f)
    assert ierr == 0
    assert (lx, ly, mx, my, nx, ny) == (LX, LY, MX, MY, NX, NY)
    assert np.array_equal(data, d)


def test_load_data_into_array(tmp_path):
    f = tmp_path / "io_test.tmp"
    d = out_data()
    io.save_data(f, "Test data", d, LX, LY, MX, MY, NX, NY)
    a = np.zeros_like(d)
    assert io.load_data_into_array(f, a, LX, LY, MX, MY, NX, NY) == 0
    assert np.array_equal(a, d)


def test_load_data_into_array_mismatch(tmp_path):
    f = tmp_path / "io_test.tmp"
    io.save_data(f, "Test data", out_data(), LX, LY, MX, MY, NX, NY)
    for b, code in (((LX, LY + 1, MX, MY, NX, NY), 1), ((LX, LY, MX + 1, MY, NX, NY), 2),
                    ((LX, LY, MX, MY + 1, NX, NY), 2), ((LX, LY, MX, MY, NX + 1, NY), 3),
                    ((LX, LY, MX, MY, NX, NY + 1), 3)):
        a = np.full(io._shape(*b), -7.0, order="F")
        assert io.load_data_into_array(f, a, *b) == code
        assert (a == -7.0).all()  # untouched


# ---------------------------------------------------------------------------
# exactness, parallel parsing, record forms

SPECIAL = [0.0, -0.0, 1.0, 0.1, 1 / 3, 5e-324, 2.2250738585072014e-308, 2.225073858507201e-308,
           1.7976931348623157e308, -1e-300, 123456789012345678.0, 16.0, -2.5e-7, 9007199254740993.0,
           np.inf, -np.inf, np.nan]


def random_bits(n, seed):
    """Finite doubles spread over the whole exponent range."""
    r = np.random.default_rng(seed)
    x = r.integers(0, 2**63 - 1, n, dtype=np.int64)
    x |= r.integers(0, 2, n, dtype=np.int64) << 63
    a = x.view(np.float64)
    return np.where(np.isfinite(a), a, 1.5)


def test_round_trip_is_bit_exact(tmp_path):
    f = tmp_path / "r.dat"
    d = np.asfortranarray(np.concatenate([SPECIAL, random_bits(40 * 7 * 9 - len(SPECIAL), 1)]).reshape(40, 7, 9, order="F"))
    io.save_data(f, "random bits", d, -1, 38, 2, 8, 0, 8)
    data, *b, ierr = io.load_data(f)
    assert ierr == 0 and tuple(b) == (-1, 38, 2, 8, 0, 8)
    assert np.array_equal(bits(data), bits(d))


def test_parallel_parse_with_blank_records_and_crlf(tmp_path):
    """A file large enough for many parsing threads, with blank records at
    random places (list-directed reads skip them), CRLF line ends and
    trailing items: the values land in order."""
    n = 600_000
    vals = random_bits(n, 2)
    r = np.random.default_rng(5)
    body = []
    for i, v in enumerate(vals):
        if r.random() < 0.01:
            body.append("   ")
        s = repr(float(v))
        body.append(f"  {s}  trailing item" if i % 997 == 0 else f" {s}")
    text = " DATA: x\r\n DIMENSIONS: Lmin Lmax Mmin Mmax Nmin Nmax\r\n 0 599 1 1000\r\n 1 1\r\n"
    text += "\r\n".join(body) + "\r\n"
    f = tmp_path / "p.dat"
    f.write_bytes(text.encode())
    assert io.data_bounds(f) == (0, 599, 1, 1000, 1, 1)
    data, *b, ierr = io.load_data(f)
    assert ierr == 0
    assert np.array_equal(bits(data.ravel(order="F")), bits(vals))


SPELLINGS = ["   1.0000000000000000     ", "  0.10000000000000001     ", "   1.0000000000000000E-002",
             "  -1.2345678901234567E+300", "   1.00000000000000000D+00", "  2.5d-3", "1.0+5", "  4.0Q1",
             "+7.25", "-.5", "5.", "  3*2.5", "1.5,2.5", "  6.02214076E23 / comment", "1e-400",
             "2.4703282292062327e-324", "2.4703282292062328e-324", "  -0.0", "Infinity", "-inf"]
SPELLED = [1.0, 0.10000000000000001, 1.0e-2, -1.2345678901234567e300, 1.0, 2.5e-3, 1.0e5, 40.0, 7.25, -0.5,
           5.0, 2.5, 1.5, 6.02214076e23, 0.0, 0.0, 5e-324, -0.0, float("inf"), float("-inf")]


def spellings_file(path, null_records=0):
    n = len(SPELLINGS) + null_records
    path.write_text(" DATA: spellings\n DIMENSIONS\n" + f" 1 {n} 1 1 1 1\n" + "\n".join(SPELLINGS) + "\n"
                    + " ,\n  /\n"[:4 * null_records])
    return n


def test_other_compilers_spellings(tmp_path):
    """Records as gfortran, ifort or older compilers write them: 17 digits
    with padding, E/D/Q exponents, exponents without a letter, leading plus,
    repeat counts, underflow (correctly rounded), null values (which leave the
    old value)."""
    f = tmp_path / "s.dat"
    n = spellings_file(f, null_records=2)
    a = np.full((n, 1, 1), 42.0, order="F")
    assert io.load_data_into_array(f, a, 1, n, 1, 1, 1, 1) == 0
    assert np.array_equal(bits(a.ravel()[:-2]), bits(SPELLED))
    assert (a.ravel()[-2:] == 42.0).all()  # null values: unchanged
    g = tmp_path / "o.dat"
    g.write_text(" DATA: o\n D\n 1 1 1 1 1 1\n 1E400\n")
    with pytest.raises(_lib.TplsError, match="not a real number"):
        io.load_data(g)  # the Fortran runtime stops on an overflowing input value


def test_bounds_may_span_records(tmp_path):
    f = tmp_path / "b.dat"
    f.write_text(" DATA: b\n DIMENSIONS\n 0 1\n 0,0\n\n 0 0 extra\n 1.5\n 2.5\n")
    data, *b, ierr = io.load_data(f)
    assert tuple(b) == (0, 1, 0, 0, 0, 0) and ierr == 0
    assert data.ravel().tolist() == [1.5, 2.5]


def test_short_and_bad_files(tmp_path):
    f = tmp_path / "short.dat"
    io.save_data(f, "short", out_data(), LX, LY, MX, MY, NX, NY)
    lines = f.read_text().splitlines(keepends=True)
    f.write_text("".join(lines[:-5]))
    data, *b, ierr = io.load_data(f)
    assert ierr == 1  # load_data's row-count status (io.F90:111-114)
    assert np.array_equal(data.ravel(order="F")[:-5], out_data().ravel(order="F")[:-5])
    a = np.zeros_like(out_data())
    with pytest.raises(_lib.TplsError, match="holds 1084 of 1089 values"):
        io.load_data_into_array(f, a, LX, LY, MX, MY, NX, NY)
    g = tmp_path / "bad.dat"
    g.write_text(" DATA: bad\n DIM\n 1 3 1 1 1 1\n 1.0\n 2.0x\n 3.0\n")
    with pytest.raises(_lib.TplsError, match="record 2 is not a real number"):
        io.load_data(g)
    h = tmp_path / "nohdr.dat"
    h.write_text(" DATA: bad\n")
    with pytest.raises(_lib.TplsError, match="ends before the bounds"):
        io.load_data(h)
    with pytest.raises(_lib.TplsError, match="No such file"):
        io.load_data(tmp_path / "missing.dat")


def test_empty_bounds(tmp_path):
    f = tmp_path / "e.dat"
    io.save_data(f, "empty", np.zeros((0, 3, 2), order="F"), 1, 0, 1, 3, 1, 2)
    data, *b, ierr = io.load_data(f)
    assert data.shape == (0, 3, 2) and ierr == 0


# ---------------------------------------------------------------------------
# Fortran list-directed I/O as the referee

needs_fortran_io = pytest.mark.skipif(not os.path.exists(DRIVE_IO), reason="tests/fortran/drive_io not built")


@needs_fortran_io
def test_fortran_written_file_reads_bit_exactly(tmp_path):
    d = np.asfortranarray(np.concatenate([SPECIAL, random_bits(11 * 5 * 6 - len(SPECIAL), 3)]).reshape(11, 5, 6, order="F"))
    raw, dat = tmp_path / "x.raw", tmp_path / "f.dat"
    d.ravel(order="F").tofile(raw)
    subprocess.run([DRIVE_IO, "write", raw, dat, "0", "10", "-2", "2", "1", "6"], check=True)
    data, *b, ierr = io.load_data(dat)
    assert tuple(b) == (0, 10, -2, 2, 1, 6) and ierr == 0
    assert np.array_equal(bits(data), bits(d))


@needs_fortran_io
def test_spellings_read_as_fortran_reads_them(tmp_path):
    f, raw = tmp_path / "s.dat", tmp_path / "s.raw"
    n = spellings_file(f)
    subprocess.run([DRIVE_IO, "read", f, raw], check=True)
    v = np.fromfile(raw, dtype=np.float64, offset=24)
    data, *b, ierr = io.load_data(f)
    assert ierr == 0 and len(v) == n
    assert np.array_equal(bits(data.ravel()), bits(v))


@needs_fortran_io
def test_library_written_file_reads_bit_exactly_in_fortran(tmp_path):
    d = np.asfortranarray(np.concatenate([SPECIAL, random_bits(11 * 5 * 6 - len(SPECIAL), 4)]).reshape(11, 5, 6, order="F"))
    dat, raw = tmp_path / "o.dat", tmp_path / "o.raw"
    io.save_data(dat, "Initial pressure", d, 0, 10, -2, 2, 1, 6)
    subprocess.run([DRIVE_IO, "read", dat, raw], check=True)
    b = np.fromfile(raw, dtype=np.int32, count=6)
    v = np.fromfile(raw, dtype=np.float64, offset=24)
    assert b.tolist() == [0, 10, -2, 2, 1, 6]
    assert np.array_equal(bits(v), bits(d.ravel(order="F")))


# ---------------------------------------------------------------------------
# options_utils_test.f90

def add_test_options(o):
    """add_test_options (options_utils_test.f90:183-195)."""
    o.put_string_option("string", "blah")
    o.put_integer_option("int", 2)
    o.put_double_option("double", 3.45)
    o.put_logical_option("logical", True)
    o.put_option("option", "value")


def add_typed_test_options(o):
    """add_typed_test_options (options_utils_test.f90:198-207)."""
    o.put_typed_option("string_type", "blah_type")
    o.put_typed_option("int_type", 6)
    o.put_typed_option("double_type", 7.89)
    o.put_typed_option("logical_type", False)


def check_test_options(o):
    assert o.get_typed_option("string", str) == ("blah", 0)
    assert o.get_typed_option("int", int) == (2, 0)
    v, ierr = o.get_typed_option("double", float)
    assert ierr == 0 and abs(v - 3.45) < 1e-12
    assert o.get_typed_option("logical", bool) == (True, 0)


def check_typed_test_options(o):
    assert o.get_typed_option("string_type", str) == ("blah_type", 0)
    assert o.get_typed_option("int_type", int) == (6, 0)
    v, ierr = o.get_typed_option("double_type", float)
    assert ierr == 0 and abs(v - 7.89) < 1e-12
    assert o.get_typed_option("logical_type", bool) == (False, 0)


def test_has_option():
    o = io.Options()
    for i in (1, 2, 3):
        o.put_option(f"option{i}", f"value{i}")
    assert o.has_option("option1") and o.has_option("option2") and o.has_option("option3")
    assert not o.has_option("option4")


def test_get_option():
    o = io.Options()
    for n, v in (("string", "blah"), ("int", "2"), ("double", "3.45"), ("logical", "true")):
        o.put_option(n, v)
    assert o.get_option("string") == ("blah", 0)
    assert o.get_string_option("string") == ("blah", 0)
    assert o.get_typed_option("string", str) == ("blah", 0)
    assert o.get_option("int") == ("2", 0)
    assert o.get_integer_option("int") == (2, 0)
    assert o.get_option("double") == ("3.45", 0)
    assert o.get_double_option("double") == (3.45, 0)
    assert o.get_option("logical") == ("true", 0)
    assert o.get_logical_option("logical") == (True, 0)


def test_get_no_such_option():
    o = io.Options()
    o.put_option("string", "blah")
    for get in (o.get_option, o.get_string_option, o.get_integer_option, o.get_double_option,
                o.get_logical_option):
        assert get("blah")[1] == io.OPTION_NOT_FOUND_ERR


def test_put_option():
    o = io.Options()
    add_test_options(o)
    add_typed_test_options(o)
    assert o.num_options == 9
    check_test_options(o)
    check_typed_test_options(o)


def test_multiple_put_option():
    o = io.Options()
    o.put_option("key", "blah")
    assert (o.num_options, o.get_option("key")) == (1, ("blah", 0))
    o.put_option("key_two", "two_blah")
    assert (o.num_options, o.get_option("key_two")) == (2, ("two_blah", 0))
    o.put_option("key", "tpls")
    assert (o.num_options, o.get_option("key")) == (2, ("tpls", 0))
    o.put_option("key", "dim")
    assert (o.num_options, o.get_option("key")) == (2, ("dim", 0))
    o.put_option("key_two", "69")
    assert (o.num_options, o.get_option("key_two")) == (2, ("69", 0))
    o.put_option("key_3", "This is another value")
    assert (o.num_options, o.get_option("key_3")) == (3, ("This is another value", 0))


HEADER = ["options_utils_test header", "Automatically created by options_utils_test", "Options now follow..."]


def test_save_and_load_options(tmp_path):
    f = tmp_path / "test_options_utils.tmp"
    o = io.Options()
    add_test_options(o)
    add_typed_test_options(o)
    o.save_options(f, HEADER)
    o2, ierr = io.load_options(f)
    assert ierr == 0 and o2.num_options == 9
    check_test_options(o2)
    check_typed_test_options(o2)
    assert f.read_text().splitlines()[:3] == ["# " + h for h in HEADER]


def test_save_append_load_options(tmp_path):
    f = tmp_path / "test_options_utils.tmp"
    o = io.Options()
    add_test_options(o)
    o.save_options(f, HEADER)
    t = io.Options()
    add_typed_test_options(t)
    t.append_options(f)
    o2, ierr = io.load_options(f)
    assert ierr == 0 and o2.num_options == 9
    check_test_options(o2)
    check_typed_test_options(o2)


def test_load_options_not_found(tmp_path):
    o, ierr = io.load_options(tmp_path / "blah.tmp")
    assert o is None and ierr == io.OPTION_FILE_NOT_FOUND_ERR


def test_load_options_format_error(tmp_path):
    f = tmp_path / "test_options_utils.tmp"
    f.write_text("# Comment\n# Comment\n# Comment\nkey value\ninteger 123\nk v\ninvalidvalue\n")
    o, ierr = io.load_options(f)
    assert ierr == io.OPTION_FORMAT_ERR
    assert o.items() == [("key", "value"), ("integer", "123"), ("k", "v")]


def test_load_duplicates(tmp_path):
    f = tmp_path / "test_options_utils.tmp"
    f.write_text("key 123\nkey 456\nkey 789\nkey abc\nkey check\n")
    o, ierr = io.load_options(f)
    assert ierr == 0 and o.num_options == 1
    assert o.get_option("key") == ("check", 0)


def test_typed_get_of_wrong_type_raises():
    o = io.Options()
    o.put_option("i", "2.5")
    o.put_option("l", "yes")
    o.put_option("d", "abc")
    o.put_option("blank", "   ")
    for get, name in ((o.get_integer_option, "i"), (o.get_logical_option, "l"), (o.get_double_option, "d"),
                      (o.get_double_option, "blank")):
        with pytest.raises(_lib.TplsError, match="is not"):
            get(name)


# ---------------------------------------------------------------------------
# the reference's options_utils.F90 as the referee

def load_report(o, ierr):
    """The library's answer in the format of ref_options_driver's "load"."""
    out = [f"ierr {ierr}", f"num {o.num_options}"]
    for n, v in o.items():
        out.append(f"{n}|{v}|")
        if n[:2] == "i_":
            x, e = o.get_integer_option(n)
            out.append(f"int {x} {e}")
        elif n[:2] == "d_":
            x, e = o.get_double_option(n)
            out.append("dbl %016X %d" % (struct.unpack("<Q", struct.pack("<d", x))[0], e))
        elif n[:2] == "l_":
            x, e = o.get_logical_option(n)
            out.append(f"log {'T' if x else 'F'} {e}")
    return "\n".join(out) + "\n"


def put_file(path, requests):
    o = io.Options()
    for line in requests.splitlines():
        if not line.strip():
            continue
        kind, rest = line[0], line[1:].lstrip()
        name, _, value = rest.partition(" ")
        if kind == "i":
            o.put_integer_option(name, int(value))
        elif kind == "d":
            o.put_double_option(name, struct.unpack("<d", struct.pack("<Q", int(value, 16)))[0])
        elif kind == "l":
            o.put_logical_option(name, value[:1] == "T")
        else:
            o.put_string_option(name, value.rstrip())
    o.save_options(path, ["ref header one", "ref header two"])


def test_load_options_matches_reference_fixture():
    o, ierr = io.load_options(os.path.join(GOLDEN, "options_battery.opt"))
    with open(os.path.join(GOLDEN, "options_battery_ref.txt")) as f:
        assert load_report(o, ierr) == f.read()


def test_put_and_save_options_match_reference_fixture(tmp_path):
    with open(os.path.join(GOLDEN, "options_put.txt")) as f:
        req = f.read()
    put_file(tmp_path / "ours.opt", req)
    with open(os.path.join(GOLDEN, "options_put_ref.opt"), "rb") as f:
        assert (tmp_path / "ours.opt").read_bytes() == f.read()


@pytest.mark.skipif(not os.path.exists(REF_OPTIONS), reason="oracle/_ref not built")
def test_options_match_live_reference(tmp_path):
    """The reference module on inputs made here: a shuffled battery with
    random names and values."""
    r = np.random.default_rng(11)
    lines = []
    for i in range(300):
        kind = r.integers(0, 4)
        pad = " " * int(r.integers(0, 4))
        if kind == 0:
            lines.append(f"{pad}i_{i} {pad}{int(r.integers(-10**9, 10**9))}")
        elif kind == 1:
            x = float(r.standard_normal() * 10.0 ** r.integers(-30, 30))
            form = r.integers(0, 3)
            s = repr(x) if form == 0 else ("%.12E" % x).replace("E", "D") if form == 1 else "%.17g" % x
            lines.append(f"{pad}d_{i} {s}")
        elif kind == 2:
            lines.append(f"l_{i} {r.choice(['T', 'F', '.true.', 'false', 't'])}")
        else:
            lines.append(f"s_{i} {'x' * int(r.integers(1, 70))} tail")
        if r.random() < 0.1:
            lines.append("# comment" if r.random() < 0.5 else "")
    f = tmp_path / "live.opt"
    f.write_text("\n".join(lines) + "\n")
    ref = subprocess.run([REF_OPTIONS, "load", f], capture_output=True, text=True, check=True).stdout
    o, ierr = io.load_options(f)
    assert load_report(o, ierr) == ref
    req = "\n".join(f"d d_{i} {struct.unpack('<Q', struct.pack('<d', float(x)))[0]:016X}"
                    for i, x in enumerate(r.standard_normal(200) * 10.0 ** r.integers(-11, 10, 200)))
    subprocess.run([REF_OPTIONS, "put", tmp_path / "ref.opt"], input=req + "\n", text=True, check=True)
    put_file(tmp_path / "ours.opt", req)
    assert (tmp_path / "ours.opt").read_bytes() == (tmp_path / "ref.opt").read_bytes()


# ---------------------------------------------------------------------------
# run configuration (main_ns_hybrid.F90:242-311, configure.F90)

def write_initial_config(d, **over):
    keys = dict(maxl=257, maxm=145, maxn=153, re=1.0, dpdl=-1.0, mu_plus=1.0, mu_minus=1.0, height=0.5,
                scap=1.0, omega_max=0.0, smooth_width=0.0098684211, dx=0.0065789474, dy=0.0065789474,
                dz=0.0065789474, dt=0.0001, idm="lsm")
    keys.update(over)
    o = io.Options()
    for k, v in keys.items():
        if v is not None:
            o.put_typed_option(k, v)
    o.save_options(os.path.join(d, "initial_config.dat"), ["TPLS initial conditions"])


def test_read_configuration_of_the_reference_sample(tmp_path):
    write_initial_config(tmp_path)
    prm, run = io.read_configuration(tmp_path, os.path.join(GOLDEN, "tpls_config.opt"))
    assert (run.maxl, run.maxm, run.maxn) == (257, 145, 153)
    assert (prm.mom_u, prm.mom_v, prm.mom_w, prm.levelset_its) == (30, 30, 30, 10)
    assert (run.num_procs_x, run.num_procs_y, run.auto_decomp) == (2, 2, 0)
    assert (run.n_timesteps, run.phi_dat_frequency, run.uvw_dat_frequency, run.backup_frequency) == (1, 1, 1, 1)
    assert (run.backup_text_format, run.backup_hdf5_format, run.idm_dim) == (1, 1, 0)
    assert run.maxu == 10.0 and run.dpdl == -1.0
    # the values exactly as the F20.10 fields hold them
    assert prm.dx == 0.0065789474 and prm.smooth_width == 0.0098684211 and prm.dt == 1e-4
    assert (prm.Re, prm.mu_plus, prm.mu_minus, prm.height, prm.scap) == (1.0, 1.0, 1.0, 0.5, 1.0)


def expect_config_error(tmp_path, ierr, message, opts=None):
    with pytest.raises(io.ConfigError) as e:
        io.read_configuration(tmp_path, opts or os.path.join(GOLDEN, "tpls_config.opt"))
    assert e.value.ierr == ierr
    assert str(e.value) == message


def test_configuration_errors(tmp_path):
    expect_config_error(tmp_path / "nodir", io.CONFIG_DIR_NOT_FOUND_ERR,
                        f"Error: Configuration directory not found: {tmp_path / 'nodir'}")
    expect_config_error(tmp_path, io.OPTION_FILE_NOT_FOUND_ERR,
                        f"Error: Options file not found: {tmp_path / 'none.opt'}", str(tmp_path / "none.opt"))
    expect_config_error(tmp_path, io.CONFIG_FILE_NOT_FOUND_ERR,
                        f"Error: Initial conditions configuration file not found: {tmp_path}/initial_config.dat")
    # the generator's input file lacks smooth_width (it holds the scale);
    # the reference reports the scale's key (configure.F90:270-274)
    import shutil
    shutil.copy(os.path.join(GOLDEN, "initial_config.opt"), tmp_path / "initial_config.dat")
    expect_config_error(tmp_path, io.OPTION_NOT_FOUND_ERR, "Error: Missing value: smooth_width_scale")
    write_initial_config(tmp_path, height=1.5)
    expect_config_error(tmp_path, io.OPTION_INVALID_ERR,
                        "Error: Invalid value: height         1.5000000000 Value must be 0 .. 1")
    write_initial_config(tmp_path, maxm=0)
    expect_config_error(tmp_path, io.OPTION_INVALID_ERR, "Error: Invalid value: maxm      0 Value must be >= 1")
    write_initial_config(tmp_path, dt=0.0)
    expect_config_error(tmp_path, io.OPTION_INVALID_ERR,
                        "Error: Invalid value: dt         0.0000000000 Value must be > 0")
    write_initial_config(tmp_path, idm="vof")
    expect_config_error(tmp_path, io.OPTION_INVALID_ERR, "Error: Invalid value: idm vof Value must lsm or dim")
    write_initial_config(tmp_path, dz=None)
    expect_config_error(tmp_path, io.OPTION_NOT_FOUND_ERR, "Error: Missing value: dz")
    write_initial_config(tmp_path, maxl=256)
    expect_config_error(tmp_path, io.OPTION_INVALID_ERR,
                        "Error: num_procs_x must divide maxl - 1.      2 does not divide    255")
    write_initial_config(tmp_path)
    bad = tmp_path / "bad.opt"
    bad.write_text(open(os.path.join(GOLDEN, "tpls_config.opt")).read().replace("num_timesteps 1", "num_timesteps 0"))
    expect_config_error(tmp_path, io.OPTION_INVALID_ERR,
                        "Error: Invalid value: num_timesteps      0 Value must be >= 1", str(bad))
    bad.write_text(open(os.path.join(GOLDEN, "tpls_config.opt")).read().replace("mom_v 30", "mom_v thirty"))
    with pytest.raises(_lib.TplsError, match="is not an integer"):
        io.read_configuration(tmp_path, bad)


def test_dim_configuration_needs_its_keys(tmp_path):
    write_initial_config(tmp_path, idm="dim")
    expect_config_error(tmp_path, io.OPTION_NOT_FOUND_ERR, "Error: Missing value: pe")
    write_initial_config(tmp_path, idm="dim", pe=1.0, epn=0.01)
    prm, run = io.read_configuration(tmp_path, os.path.join(GOLDEN, "tpls_config.opt"))
    assert run.idm_dim == 1 and run.max_iteration_dim == 18 and run.epn == 0.01


def test_written_initial_conditions_read_back(tmp_path):
    from tpls_amd.config import CaseParams
    from tpls_amd.initial import make_initial
    p = CaseParams(nx=16, ny=8, nz=12, Re=50.0, mu_plus=3.0, dt=2e-4, seed=4)
    fields = make_initial(p)
    io.write_initial_conditions(tmp_path, p, fields)
    io.write_options_file(tmp_path / "tpls_config.opt", p, n_timesteps=7)
    prm, run = io.read_configuration(tmp_path, tmp_path / "tpls_config.opt")
    assert (run.maxl, run.maxm, run.maxn, run.n_timesteps) == (p.maxl, p.maxm, p.maxn, 7)
    assert prm.Re == 50.0 and prm.mu_plus == 3.0 and prm.dt == 2e-4
    assert abs(prm.dx - p.dx) <= 5e-11 and abs(prm.smooth_width - p.smooth_width) <= 5e-11
    for a, name, up in zip(fields, ("u", "v", "w", "pressure", "phi"),
                           (p.maxn - 2, p.maxn - 2, p.maxn - 1, p.maxn, p.maxn)):
        b = np.zeros_like(a)
        assert io.load_data_into_array(tmp_path / io.INITIAL_FILES[name], b, 0, p.maxl, 0, p.maxm, 0, up) == 0
        assert np.array_equal(bits(a), bits(b))


def test_written_dim_configuration_reads_back_into_the_case_params(tmp_path):
    """An idm dim case written in the generator's layout (pe and epn after
    idm, create_initial_conditions_program.F90:152-156; max_iteration_dim in
    the options) reads back into tpls_params' idm_dim, epn, Pe, dim_its, and
    the phase field is the generator's tanh profile in [0, 1]."""
    from tpls_amd.config import CaseParams
    from tpls_amd.initial import make_initial
    p = CaseParams(nx=16, ny=8, nz=16, seed=2, idm="dim", dim_its=11)
    fields = make_initial(p)
    assert 0.0 < fields[4].min() and fields[4].max() < 1.0
    io.write_initial_conditions(tmp_path, p, fields)
    io.write_options_file(tmp_path / "tpls_config.opt", p, n_timesteps=2)
    keys = [l.split()[0] for l in open(tmp_path / io.INITIAL_CONFIG_FILE) if l.strip() and not l.startswith("#")]
    assert keys[-3:] == ["idm", "pe", "epn"]
    prm, run = io.read_configuration(tmp_path, tmp_path / "tpls_config.opt")
    assert run.idm_dim == 1 and prm.idm_dim == 1 and prm.dim_its == run.max_iteration_dim == 11
    # F20.10: epn = dz/2 = 1/32 and Pe = 1024 are exact
    assert prm.epn == p.dim_epn == 1 / 32 and prm.Pe == p.dim_Pe == 1024.0
    lsm = CaseParams(nx=16, ny=8, nz=16, seed=2)
    io.write_initial_conditions(tmp_path, lsm, make_initial(lsm))
    io.write_options_file(tmp_path / "tpls_config.opt", lsm, n_timesteps=2)
    prm, run = io.read_configuration(tmp_path, tmp_path / "tpls_config.opt")
    assert prm.idm_dim == 0 and prm.dim_its == 0


# ---------------------------------------------------------------------------
# channel outputs and the text backup (io.F90:171-791)

def grid_fields(l, m, n, seed):
    r = np.random.default_rng(seed)
    planes = {"uv": n - 1, "w": n, "p": n + 1}
    return {k: np.asfortranarray(r.standard_normal((l + 1, m + 1, planes[g]))) for k, g in io.BACKUP_ARRAYS}


def records(path, skip):
    return np.loadtxt(path, skiprows=skip, ndmin=2)


def test_channel_output_phi_and_input(tmp_path):
    l, m, n, dx, dy, dz = 7, 5, 6, 0.125, 0.1, 1 / 3
    phi = grid_fields(l, m, n, 1)["phi"]
    f = tmp_path / "phi_channel.dat"
    io.channel_output_phi(phi, l, m, n, dx, dy, dz, f)
    head = f.read_text().splitlines()[:2]
    assert head[0] == ' VARIABLES="X ","Y ","Z ","Phi"'
    assert head[1].split() == ['ZONE', 'T="Floor",', 'I=', str(l - 1), 'J=', str(m - 1), 'K=', str(n - 1), 'F=POINT']
    rec = records(f, 2)
    assert rec.shape == ((l - 1) * (m - 1) * (n - 1), 4)
    # records: k slowest, then j, then i = 1..l-1; cell-centre coordinates
    i, j, k = np.meshgrid(np.arange(1, l), np.arange(0, m - 1), np.arange(0, n - 1), indexing="ij")
    i, j, k = (a.ravel(order="F") for a in (i, j, k))
    assert np.array_equal(rec[:, 0], dx / 2 + dx * i) and np.array_equal(rec[:, 2], dz / 2 + dz * k)
    assert np.array_equal(rec[:, 1], dy / 2 + dy * j)
    assert np.array_equal(rec[:, 3], phi[i + 1, j, k + 1])
    back = np.zeros_like(phi)
    assert io.channel_input_phi(f, back, l, m, n) == 0
    assert np.array_equal(back[2:, :m - 1, 1:n], phi[2:, :m - 1, 1:n])
    assert (back[:2] == 0).all() and (back[:, m - 1:] == 0).all()


def test_channel_output_uvw_and_2d(tmp_path):
    l, m, n, dx, dy, dz = 6, 7, 5, 0.2, 0.25, 0.5
    a = grid_fields(l, m, n, 2)
    u3, v3, w3, p, phi, visc = a["u2"], a["v2"], a["w2"], a["pres"], a["phi"], a["pres_old"]
    f = tmp_path / "uvw.dat"
    io.channel_output_uvw(u3, v3, w3, p, l, m, n, dx, dy, dz, f)
    rec = records(f, 2)
    i, j, k = np.meshgrid(np.arange(1, l), np.arange(0, m - 1), np.arange(0, n - 1), indexing="ij")
    i, j, k = (x.ravel(order="F") for x in (i, j, k))
    # face averages to the cell centre, pressure one plane up (io.F90:575-581)
    assert np.array_equal(rec[:, 3], (u3[i, j, k] + u3[i + 1, j, k]) / 2)
    assert np.array_equal(rec[:, 4], (v3[i, j, k] + v3[i, j + 1, k]) / 2)
    assert np.array_equal(rec[:, 5], (w3[i, j, k] + w3[i, j, k + 1]) / 2)
    assert np.array_equal(rec[:, 6], p[i, j, k + 1])
    g = tmp_path / "2d.dat"
    io.channel_output2d(u3, v3, w3, p, phi, visc, l, m, n, dx, dy, dz, g)
    assert g.read_text().splitlines()[1].split()[4:6] == ["K=", str(n - 1)]
    rec = records(g, 2)
    i, k = np.meshgrid(np.arange(1, l), np.arange(0, n - 1), indexing="ij")
    i, k = i.ravel(order="F"), k.ravel(order="F")
    jm = (m - 1) // 2
    assert rec.shape == ((l - 1) * (n - 1), 8)
    assert np.array_equal(rec[:, 1], dz / 2 + dz * k)
    assert np.array_equal(rec[:, 6], phi[i + 1, jm, k + 1]) and np.array_equal(rec[:, 7], visc[i, jm, k + 1])


def test_backup_channel_round_trip_and_layout(tmp_path):
    l, m, n = 5, 4, 6
    a = grid_fields(l, m, n, 3)
    f = tmp_path / "fieldbackup1.dat"
    io.backup_channel(a, l, m, n, f)
    lines = f.read_text().splitlines()
    plane = (l + 1) * (m + 1)
    assert len(lines) == plane * (2 * (n - 1) + n + 2 * (n + 1))
    first = [float(x) for x in lines[0].split()]
    assert first == [a["u2"][0, 0, 0], a["conv0_u"][0, 0, 0], a["conv1_u"][0, 0, 0]]
    last = [float(x) for x in lines[-1].split()]
    assert last == [a["pres"][l, m, n], a["pres_old"][l, m, n]]
    b = io.dataload_channel(f, l, m, n)
    for k in a:
        assert np.array_equal(b[k].view(np.int64), a[k].view(np.int64)), k
    f.write_text("\n".join(lines[: len(lines) // 2]) + "\n")
    with pytest.raises(_lib.TplsError, match="ends inside a backup block"):
        io.dataload_channel(f, l, m, n)


def test_gathered_ghost_rules():
    from tpls_amd.run import gathered
    maxl, maxm, maxn = 6, 5, 7
    r = np.random.default_rng(4)
    a = np.asfortranarray(r.standard_normal((maxl + 1, maxm + 1, maxn - 1)))
    inlet = r.standard_normal(maxn - 1)
    u = gathered("u", a, maxl, maxm, maxn, u_inlet=inlet)
    assert (u[0] == inlet[None, :]).all()
    assert np.array_equal(u[maxl], u[maxl - 2]) and np.array_equal(u[maxl - 1, 1:maxm], a[maxl - 2, 1:maxm])
    assert np.array_equal(u[:, 0], u[:, maxm - 1]) and np.array_equal(u[1:maxl - 1, 1:maxm], a[1:maxl - 1, 1:maxm])
    w = gathered("w", np.asfortranarray(r.standard_normal((maxl + 1, maxm + 1, maxn))), maxl, maxm, maxn)
    assert (w[:, :, 0] == 0).all() and (w[:, :, maxn - 1] == 0).all() and (w[0] == 0).all() and (w[maxl] == 0).all()
    phi = gathered("phi", np.asfortranarray(r.standard_normal((maxl + 1, maxm + 1, maxn + 1))), maxl, maxm, maxn,
                   dz=0.25, height=0.5)
    assert phi[0, 2, 3] == 3 * 0.25 - 0.5 * 0.25 - 0.5
    assert np.array_equal(phi[:, :, maxn], phi[:, :, maxn - 1]) and np.array_equal(phi[:, :, 0], phi[:, :, 1])
