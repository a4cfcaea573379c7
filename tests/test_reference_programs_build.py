"""The reference's own test programs (test/ib_test.cc, ib_ns_test.cc,
mg_test.cc, mg_test2.cc, mg_test3.cc, ns_test.cc, restr_test.cc) compile and
link unmodified against the drop-in headers (`-I include`, the reference's
include paths dt/, alg/, mg/, ns/, ib/) and libafm.so.

Compile and link only: the programs are read where they lie in the reference
tree and never copied; the binaries go to a temporary directory and are not
run (this test has no GPU; the same API is run on the GPU by
tests/cpp/dropin_main.cc).  Skipped where the reference tree is absent.
AFM_REFERENCE_DIR names the tree (default /root/reference)."""
import concurrent.futures as cf
import glob
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("AFM_REFERENCE_DIR", "/root/reference")
PROGRAMS = ["ib_ns_test", "ib_test", "mg_test", "mg_test2", "mg_test3", "ns_test", "restr_test"]
LIB = os.path.join(ROOT, "afmibins_amd", "lib", "libafm.so")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "test")), reason="reference tree absent")


def test_reference_program_set_is_the_expected_one():
    found = sorted(os.path.basename(p)[:-3] for p in glob.glob(os.path.join(REF, "test", "*.cc")))
    assert found == sorted(PROGRAMS)


@pytest.mark.parametrize("std", ["c++14"])
def test_reference_programs_compile_and_link(tmp_path, std):
    link = os.path.exists(LIB)

    def build(name):
        src = os.path.join(REF, "test", name + ".cc")
        cmd = ["g++", "-std=" + std, "-O0", "-I", os.path.join(ROOT, "include"), src]
        if link:
            cmd += ["-L", os.path.dirname(LIB), "-lafm", "-o", str(tmp_path / name)]
        else:
            cmd += ["-c", "-o", str(tmp_path / (name + ".o"))]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        return name, r.returncode, r.stderr

    with cf.ThreadPoolExecutor(4) as pool:
        results = list(pool.map(build, PROGRAMS))
    failed = [(n, err[-2000:]) for n, rc, err in results if rc != 0]
    assert not failed, failed
    for name in PROGRAMS:
        assert (tmp_path / (name if link else name + ".o")).exists()


def test_headers_alone_compile_under_each_standard(tmp_path):
    """Every header at the reference's paths compiles on its own (it includes
    what it uses) under C++11, 14 and 17."""
    headers = sorted(glob.glob(os.path.join(ROOT, "include", d, "*.h")) for d in ("dt", "alg", "mg", "ns", "ib"))
    headers = [h for group in headers for h in group]
    assert len(headers) >= 17

    def check(args):
        std, h = args
        src = tmp_path / (std.replace("+", "p") + "_" + os.path.basename(h).replace(".", "_") + ".cc")
        rel = os.path.relpath(h, os.path.join(ROOT, "include"))
        src.write_text('#include "%s"\nint main() { return 0; }\n' % rel)
        r = subprocess.run(["g++", "-std=" + std, "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                           capture_output=True, text=True, timeout=300)
        return rel, std, r.returncode, r.stderr

    with cf.ThreadPoolExecutor(4) as pool:
        results = list(pool.map(check, [(s, h) for s in ("c++11", "c++14", "c++17") for h in headers]))
    failed = [(rel, std, err[-1500:]) for rel, std, rc, err in results if rc != 0]
    assert not failed, failed
