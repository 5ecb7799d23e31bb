"""Host-side failure handling (CPU, no device): malformed profiles, attenuation tables and
dispersion files, invalid option combinations and unwritable outputs end in the reference's
message and a status code, before any device work; and the same host code built with g++
under AddressSanitizer + UndefinedBehaviorSanitizer survives every malformed input.

Reference error paths: Atmosphere1D.cpp:85-285 (header / data parsing), ESSModeSolver.cpp:
98-114 (f_min < f_max), 176-210 (missing vectors, thrown), ModalBroadbandPropagator.cpp:
122-125 (dispersion format error, exit), ModeSolver.cpp:319-323 (outputs opened unchecked)."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ncpaprop-release_amd")
BIN = os.path.join(PKG, "bin")

GOOD_HEADER = "#% 0, Z0, m, 0.0\n#% 1, Z, km\n#% 2, U, m/s\n#% 3, V, m/s\n#% 4, T, degK\n#% 5, RHO, g/cm3\n#% 6, P, mbar\n"
GOOD_ROWS = "".join("%g %g 0 %g %g %g\n" % (z, 5.0, 288 - 6.5 * z, 1.2e-3 * (1 - 0.09 * z), 1013 * (1 - 0.11 * z))
                    for z in (0.0, 1.0, 2.0, 3.0, 4.0))

# (name, file text, message the check must give)
PROFILE_CASES = [
    ("no_header", GOOD_ROWS, "No descriptive fields found"),
    ("short_header", "#% 1, Z\n" + GOOD_ROWS, "Error parsing descriptive line"),
    ("bad_column", "#% x, Z, km\n#% 2, U, m/s\n0 1\n1 1\n2 1\n", "First field not parseable as an integer"),
    ("negative_column", "#% -3, Z, km\n#% 2, U, m/s\n0 1\n1 1\n2 1\n", "column number out of range"),
    ("bad_units", GOOD_HEADER.replace("degK", "furlongs"), "Unrecognized units string"),
    ("duplicate_key", GOOD_HEADER + "#% 7, U, m/s\n" + "".join(l + " 1\n" for l in GOOD_ROWS.splitlines()),
     "key values are not unique"),
    ("missing_column", GOOD_HEADER + GOOD_ROWS + "5 5 0 255\n", "6 columns expected, 4 columns found"),
    ("not_a_number", GOOD_HEADER + GOOD_ROWS.replace("288", "x288", 1), "Can't parse field x288 as a double"),
    ("out_of_range", GOOD_HEADER + GOOD_ROWS.replace("288", "1e999", 1), "is out of range for double precision"),
    ("too_few_rows", GOOD_HEADER + "".join(GOOD_ROWS.splitlines(True)[:2]), "need at least 3 data rows"),
    ("missing_U", GOOD_HEADER.replace("2, U,", "2, UU,"), "U wind vector not found"),
    ("missing_V", GOOD_HEADER.replace("3, V,", "3, VV,"), "V wind vector not found"),
    ("no_sound_speed", "#% 1, Z, km\n#% 2, U, m/s\n#% 3, V, m/s\n0 1 1\n1 1 1\n2 1 1\n",
     "Can't calculate static sound speed"),
    ("missing_T", GOOD_HEADER.replace("4, T,", "4, TT,"), "T temperature vector not found"),
    ("missing_P", "#% 1, Z, km\n#% 2, U, m/s\n#% 3, V, m/s\n#% 4, T, K\n#% 5, RHO, g/cm3\n"
                  "0 1 1 288 1e-3\n1 1 1 280 1e-3\n2 1 1 270 1e-3\n", "P pressure vector not found"),
    ("non_increasing", GOOD_HEADER + GOOD_ROWS.replace("2 5 0", "0.5 5 0", 1), "altitudes must increase"),
    ("nan_value", GOOD_HEADER + GOOD_ROWS.replace("288", "nan", 1), "non-finite"),
]
for i, (name, text, msg) in enumerate(PROFILE_CASES):
    if name in ("missing_U", "missing_V", "missing_T", "bad_units"):
        PROFILE_CASES[i] = (name, text + GOOD_ROWS, msg)


def _lib():
    from ncpamodes import _lib as L
    return L.load()


def test_profile_check_accepts_good_profile(tmp_path):
    from ncpamodes import _lib as L
    p = tmp_path / "good.dat"
    p.write_text(GOOD_HEADER + GOOD_ROWS)
    info = L.ProfileInfo()
    msg = C.create_string_buffer(512)
    assert _lib().ncpa_profile_check(str(p).encode(), b"", C.byref(info), msg, 512) == 0
    assert info.n_nodes == 5 and info.has_z0 == 1 and info.z_top == pytest.approx(4000.0)


@pytest.mark.parametrize("name,text,expect", PROFILE_CASES, ids=[c[0] for c in PROFILE_CASES])
def test_profile_check_rejects(tmp_path, name, text, expect):
    from ncpamodes import _lib as L
    p = tmp_path / (name + ".dat")
    p.write_text(text)
    msg = C.create_string_buffer(512)
    rc = _lib().ncpa_profile_check(str(p).encode(), None, None, msg, 512)
    assert rc == L.NCPA_ERR_ARG
    assert expect in msg.value.decode(), msg.value


def test_profile_check_missing_file():
    msg = C.create_string_buffer(512)
    assert _lib().ncpa_profile_check(b"/nonexistent/profile.dat", None, None, msg, 512) != 0
    assert "Input stream not in good state" in msg.value.decode()


def _run(prog, args, cwd):
    return subprocess.run([os.path.join(BIN, prog)] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


def _good_profile(tmp_path):
    p = tmp_path / "good.dat"
    p.write_text(GOOD_HEADER + GOOD_ROWS)
    return str(p)


def test_modess_malformed_profile_fails_before_device(tmp_path):
    """The message of the profile reader and exit status 1, on a host without a device too."""
    p = tmp_path / "bad.dat"
    p.write_text(GOOD_HEADER + GOOD_ROWS + "5 5 0 255\n")
    out = _run("Modess", ["--atmosfile", str(p), "--freq", "0.1", "--singleprop", "--azimuth", "90"], tmp_path)
    assert out.returncode == 1
    assert "6 columns expected, 4 columns found" in out.stderr
    assert "device context" not in out.stderr


def test_modess_unwritable_output(tmp_path):
    out = _run("Modess", ["--atmosfile", _good_profile(tmp_path), "--freq", "0.1", "--singleprop", "--azimuth", "90",
                          "--filetag", "no_such_dir/run"], tmp_path)
    assert out.returncode == 1
    assert "cannot open no_such_dir/run.tloss_1d.nm for writing" in out.stderr
    ro = tmp_path / "ro"
    ro.mkdir()
    os.chmod(ro, 0o500)
    try:
        if os.access(ro, os.W_OK):
            pytest.skip("running as a user that ignores directory permissions")
        out = _run("Modess", ["--atmosfile", _good_profile(tmp_path), "--freq", "0.1", "--singleprop", "--azimuth",
                              "90", "--write_2d_tloss", "--filetag", "ro/x"], tmp_path)
        assert out.returncode == 1 and "cannot open ro/x.tloss_1d.nm" in out.stderr
    finally:
        os.chmod(ro, 0o700)


def test_modess_malformed_attenuation_file(tmp_path):
    a = tmp_path / "attn.dat"
    a.write_text("10 1e-6\n20 2e-6 7\n")
    out = _run("Modess", ["--atmosfile", _good_profile(tmp_path), "--freq", "0.1", "--singleprop", "--azimuth", "90",
                          "--use_attn_file", str(a)], tmp_path)
    assert out.returncode == 1 and "Error parsing attenuation line" in out.stderr
    a.write_text("20 1e-6\n10 2e-6\n")
    out = _run("Modess", ["--atmosfile", _good_profile(tmp_path), "--freq", "0.1", "--singleprop", "--azimuth", "90",
                          "--use_attn_file", str(a)], tmp_path)
    assert out.returncode == 1 and "altitudes must increase" in out.stderr


def test_modbb_fmin_not_below_fmax(tmp_path):
    """ESSModeSolver.cpp:101-103 throws "f_min must be less than f_max!"."""
    out = _run("ModBB", ["--dispersion", "--dispersion_file", "d.dat", "--atmosfile", _good_profile(tmp_path),
                         "--azimuth", "90", "--f_min", "0.5", "--f_step", "0.01", "--f_max", "0.5", "--method",
                         "modess"], tmp_path)
    assert out.returncode == 1 and "f_min must be less than f_max!" in out.stderr


def test_modbb_ranges_and_dispersion_file(tmp_path):
    d = tmp_path / "disp.dat"
    d.write_text("1.0e-02   2    1.2e-03   1.2e-03   1.8e-02   1.0e-08   1.0e-03   1.0e-03   1.9e-02\n")
    base = ["--propagation", "--input_dispersion_file", str(d), "--output_waveform_file", "w.dat", "--source",
            "impulse"]
    out = _run("ModBB", base + ["--receiver", "multiple", "--start_range_km", "300", "--end_range_km", "200",
                                "--range_step_km", "10"], tmp_path)
    assert out.returncode == 1 and "end_range_km must be greater than start_range_km" in out.stderr
    out = _run("ModBB", base + ["--receiver", "single", "--range_km", "200"], tmp_path)
    assert out.returncode == 1 and "read_dispersion_data format error" in out.stderr
    out = _run("ModBB", ["--propagation", "--input_dispersion_file", str(tmp_path / "missing.dat"),
                         "--output_waveform_file", "w.dat", "--receiver", "single", "--range_km", "200"], tmp_path)
    assert out.returncode == 1 and "missing.dat could not be opened." in out.stderr


# ------------------------------------------------------------------------------------------
# the host code under AddressSanitizer + UndefinedBehaviorSanitizer
# ------------------------------------------------------------------------------------------
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


@pytest.fixture(scope="module")
def host_inputs(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("san") / "host_inputs")
    srcs = [os.path.join(ROOT, "tests", "sanitize", "host_inputs.cpp"), os.path.join(PKG, "host", "profile_parse.cpp"),
            os.path.join(PKG, "host", "dispersion.cpp"), os.path.join(PKG, "cli", "params.cpp")]
    cmd = ["g++", "-std=c++17", "-pthread", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "cli")] + \
        SAN_FLAGS + srcs + ["-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + r.stderr[-3000:])
    with open(exe, "rb") as fh:
        blob = fh.read()
    assert b"__asan_init" in blob and b"__ubsan_handle" in blob   # the instrumented build
    return exe


def _san(exe, args, cwd):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    return r.stdout


def test_sanitized_profile_reader(host_inputs, tmp_path):
    files = []
    for name, text, _ in PROFILE_CASES:
        p = tmp_path / (name + ".dat")
        p.write_text(text)
        files.append(str(p))
    # degenerate texts: empty, binary junk, a header line of 4096 characters, long numbers
    (tmp_path / "empty.dat").write_text("")
    (tmp_path / "junk.dat").write_bytes(bytes(range(256)) * 8)
    (tmp_path / "long.dat").write_text("#% 1, " + "Z" * 4096 + ", km\n" + "1" * 5000 + "\n")
    (tmp_path / "huge.dat").write_text(GOOD_HEADER + GOOD_ROWS.replace("288", "1e999", 1))
    p = tmp_path / "good.dat"
    p.write_text(GOOD_HEADER + GOOD_ROWS)
    files += [str(tmp_path / n) for n in ("empty.dat", "junk.dat", "long.dat", "huge.dat", "good.dat", "missing.dat")]
    out = _san(host_inputs, ["profile"] + files, tmp_path)
    lines = dict(l.split(": ", 1) for l in out.splitlines())
    assert lines[str(p)].startswith("0 ok")
    assert all(not v.startswith("0 ") for k, v in lines.items() if k != str(p))


def test_sanitized_dispersion_reader(host_inputs, tmp_path):
    good = "1.0e-02   1    1.2e-03   1.2e-03   1.8e-02   1.0e-08   1.0e-03   1.0e-03\n" \
           "2.0e-02   2    1.2e-03   1.2e-03   3.6e-02   2.0e-08   1.0e-03   1.0e-03   3.5e-02   2.0e-08   1.0e-03   1.0e-03\n"
    cases = {"good": good, "truncated": good[:-40], "count_too_big": "1e-2 99999 1 1 1 1 1 1\n",
             "negative_count": "1e-2 -3 1 1\n", "fractional_count": "1e-2 1.5 1 1 1 1 1 1\n",
             "text": "frequency modes\n", "empty": "", "nan": "nan 1 1 1 1 1 1 1\n"}
    files = []
    for k, v in cases.items():
        (tmp_path / (k + ".dat")).write_text(v)
        files.append(str(tmp_path / (k + ".dat")))
    out = _san(host_inputs, ["dispersion"] + files, tmp_path)
    res = dict(l.split(": ", 1) for l in out.splitlines())
    assert res[files[0]].startswith("0 2 records 3 modes")
    for f in files[1:4]:
        assert not res[f].startswith("0 "), (f, res[f])


def test_sanitized_option_parser(host_inputs, tmp_path):
    (tmp_path / "modess.param").write_text("# comment\nfreq = 0.1\n:::\n=\nazimuth\n" + "x" * 3000 + "\n")
    for args in (["--freq"], ["--freq", "--azimuth"], ["-"], ["--"], ["---x"], ["--freq=abc"], ["--Nz_grid", "1e99"],
                 ["--azimuth", "nan", "--singleprop", "--atmosfile", "a", "--freq", "0.1"], ["-abc"], ["--bogus", "1"],
                 ["--singleprop", "--multiprop", "--atmosfile", "a", "--freq", "1"]):
        out = _san(host_inputs, ["options"] + args, tmp_path)
        assert out.startswith("options: "), (args, out)
