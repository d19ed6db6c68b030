"""Fuzzing of the NC.inp parser (sonium_amd/csrc/numcalc.cpp), the one place
where the library reads untrusted text.  The reference fuzzes its own config
parser for panics (crates/sonium-bem/fuzz/fuzz_targets/config_parser.rs);
here the parser must return a status (OK or a typed error with a message) for
every input:

* in process, through the C ABI of libsbh.so: arbitrary byte strings
  (hypothesis, derandomized) and seeded mutations of the reference's fixture
  projects;
* the same mutated corpus through a host-only build of the parser under
  AddressSanitizer and UndefinedBehaviorSanitizer (tests/fuzz/nc_parse_driver.cpp),
  which turns an out-of-bounds access, use after free, signed overflow or
  invalid shift into a failed run.
"""

import ctypes
import os
import subprocess

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

import sonium_amd as sa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURES = [os.path.join(ROOT, "tests", "golden", "numcalc", d, "NC.inp") for d in ("mesh2hrtf_basic", "sample_project")]
TOKENS = ["0", "-1", "1e308", "-1e308", "nan", "inf", "-inf", "1e", "0x10", "99999999999999999999", "2147483648",
          "-2147483649", "9223372036854775808", "1.5", "", " ", "\t", "ELEM", "VELO", "PRES", "ADMI", "##", "#",
          "BOUNDARY", "END", "RETURN", "FREQUENCYCURVE", "POINT SOURCES", "PLANE WAVES", "NODES", "ELEMENTS",
          "é", "\x00", "\x7f", "-0", "+", "1.0e-400", "18446744073709551616"]


def _mutants(seed=20261203, count=400):
    rng = np.random.default_rng(seed)
    bases = [open(p, encoding="utf-8", errors="replace").read() for p in FIXTURES]
    out = []
    for i in range(count):
        lines = bases[i % len(bases)].split("\n")
        for _ in range(int(rng.integers(1, 6))):
            op = int(rng.integers(0, 6))
            j = int(rng.integers(0, len(lines)))
            if op == 0 and len(lines) > 1:
                del lines[j]
            elif op == 1:
                lines.insert(j, lines[int(rng.integers(0, len(lines)))])
            elif op == 2:
                words = lines[j].split(" ")
                words[int(rng.integers(0, len(words)))] = TOKENS[int(rng.integers(0, len(TOKENS)))]
                lines[j] = " ".join(words)
            elif op == 3:
                lines[j] = lines[j][: int(rng.integers(0, len(lines[j]) + 1))]
            elif op == 4:
                lines.insert(j, " ".join(TOKENS[int(rng.integers(0, len(TOKENS)))] for _ in range(int(rng.integers(1, 8)))))
            else:
                lines = lines[: j + 1]
        out.append("\n".join(lines).replace("\x00", ""))
    for i in range(100):
        b = rng.integers(1, 256, size=int(rng.integers(0, 2048)), dtype=np.uint8).tobytes()
        out.append(b.decode("latin-1"))
    return out


def _parse(text):
    """sbh_nc_parse through the C ABI: OK (the project is destroyed again) or a
    typed error (1 invalid input, 2 unsupported keyword) with a message."""
    h = ctypes.c_void_p()
    st = sa._lib().sbh_nc_parse(text.encode("utf-8", errors="replace"), b".", ctypes.byref(h))
    if st == 0:
        sa._lib().sbh_nc_project_destroy(h)
        return 0
    assert st in (1, 2), st
    msg = sa._lib().sbh_last_error()
    assert msg and msg.decode(errors="replace").strip(), "error without a message"
    return st


def test_fixture_mutations_return_a_status():
    statuses = [_parse(t) for t in _mutants()]
    assert len(statuses) == 500
    # the corpus exercises both outcomes
    assert statuses.count(0) > 0 and sum(1 for s in statuses if s != 0) > 100


@settings(max_examples=300, derandomize=True, deadline=None)
@given(st.binary(max_size=4096))
def test_arbitrary_bytes_return_a_status(data):
    _parse(data.decode("utf-8", errors="replace").replace("\x00", ""))


def test_mutations_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "nc_fuzz"
    cmd = [hipcc, "-x", "c++", "-std=c++17", "-O1", "-g", "-ffp-contract=off",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=undefined",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "sonium_amd", "csrc"),
           "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
           os.path.join(ROOT, "sonium_amd", "csrc", "numcalc.cpp"), os.path.join(ROOT, "tests", "fuzz", "nc_parse_driver.cpp"),
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    files = []
    for i, t in enumerate(_mutants(seed=7)):
        p = tmp_path / f"case{i:04d}.inp"
        p.write_bytes(t.encode("utf-8", errors="replace"))
        files.append(str(p))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)] + files, capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout.strip())
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-3000:])
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
