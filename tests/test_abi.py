"""The C-ABI library: it loads, exports every entry point include/qcbc.h
declares, and refuses to compute without a gfx950 GPU (no CPU fallback)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qcbc.h")
LIB = os.path.join(ROOT, "quickcbc_amd", "libqcbc.so")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", txt):
        name = m.group(1)
        if name in ("if", "while", "for", "sizeof", "return"):
            continue
        names.add(name)
    return sorted(names)


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_header_declares_the_reference_entry_points():
    names = declared_functions()
    for ref in ("log_likelihood_full", "fulltemplates", "IMRPhenomDGenerateh22FDAmpPhase", "fourier_nwip",
                "templates", "log_likelihood_intrinsic",
                "F_ant", "Times", "fbegin", "CreateAmpPhaseFDWaveform", "DestroyAmpPhaseFDWaveform",
                "fullphaseamp", "heterodyne", "log_likelihood_het", "Fisher_Het", "FisherEvec"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libqcbc.so not built"
    missing = [n for n in declared_functions() if n not in exported_symbols()]
    assert not missing, missing


def test_library_loads_and_reports_errors():
    L = ctypes.CDLL(LIB)
    L.qcbc_strerror.restype = ctypes.c_char_p
    assert L.qcbc_strerror(33) == b"parameter point outside the IMRPhenomD domain"


def test_no_cpu_fallback_without_gpu():
    """Without a usable gfx950 device the batch API reports QCBC_ENODEV
    instead of computing anything on the host."""
    code = (
        "import ctypes,sys\n"
        "sys.path.insert(0, %r)\n"
        "from quickcbc_amd import qcbc as Q, workloads as W\n"
        "import ctypes\n"
        "h = ctypes.c_void_p()\n"
        "d = Q.make_desc(W.config0())\n"
        "print(Q.lib().qcbc_ctx_create(ctypes.byref(d), ctypes.byref(h)))\n" % ROOT)
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = "-1"
    env["ROCR_VISIBLE_DEVICES"] = "-1"
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().splitlines()[-1] == "101"  # QCBC_ENODEV


def test_oracle_is_not_linked_into_the_product():
    out = subprocess.run(["nm", "-D", LIB], capture_output=True, text=True, check=True).stdout
    assert "qo_" not in out
    deps = subprocess.run(["ldd", LIB], capture_output=True, text=True).stdout
    assert "oracle" not in deps


def test_point_record_mirror_matches_the_struct(tmp_path):
    """Every field of the Python mirror of qcbc::Point sits at the offset g++
    gives it, and the sizes agree."""
    from quickcbc_amd.qcbc import PointRecord
    names = [f[0] for f in PointRecord._fields_]
    src = ['#include <cstdio>', '#include <cstddef>', '#include "qcbc_types.h"', "int main() {"]
    for n in names:
        src.append('  std::printf("%%s %%zu\\n", "%s", offsetof(qcbc::Point, %s));' % (n, "lambda" if n == "lambda_" else n))
    src.append('  std::printf("sizeof %zu\\n", sizeof(qcbc::Point));')
    src.append("}")
    cpp, exe = tmp_path / "off.cpp", tmp_path / "off"
    cpp.write_text("\n".join(src) + "\n")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "quickcbc_amd", "csrc"), "-o", str(exe), str(cpp)],
                   check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True)
               .stdout.splitlines())
    assert int(got.pop("sizeof")) == ctypes.sizeof(PointRecord)
    for n in names:
        assert getattr(PointRecord, n).offset == int(got[n]), n


def test_reference_struct_mirrors_match_the_header(tmp_path):
    """struct Net and struct Het (QuickCBC.h:13-50), qcbc_desc and
    qcbc_update_desc as the ctypes bindings lay them out are the header's,
    field by field."""
    from quickcbc_amd.qcbc import Desc, HetStruct, Net, UpdateDesc
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "qcbc.h"', "int main(void) {"]
    types = (("Net", "struct Net", Net), ("Het", "struct Het", HetStruct), ("Desc", "qcbc_desc", Desc),
             ("UpdateDesc", "qcbc_update_desc", UpdateDesc))
    for tag, ctype, cls in types:
        for f in cls._fields_:
            src.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (tag, f[0], ctype, f[0]))
        src.append('  printf("%s.sizeof %%zu\\n", sizeof(%s));' % (tag, ctype))
    src.append("  return 0;\n}")
    c, exe = tmp_path / "off.c", tmp_path / "off"
    c.write_text("\n".join(src) + "\n")
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(c)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True)
               .stdout.splitlines())
    for tag, _, cls in types:
        assert int(got[tag + ".sizeof"]) == ctypes.sizeof(cls), tag
        for f in cls._fields_:
            assert getattr(cls, f[0]).offset == int(got[tag + "." + f[0]]), (tag, f[0])


def _dropin_with_freq(shift):
    """Calls log_likelihood_full in a child process with the analysis grid
    f_i = i/Tobs, or with that grid shifted by `shift` Hz from bin 2 on, and
    returns the process's exit code and stderr."""
    code = (
        "import ctypes,sys\n"
        "sys.path.insert(0, %r)\n"
        "import numpy as np\n"
        "from quickcbc_amd import qcbc as Q, workloads as W\n"
        "an = W.config0()\n"
        "net = Q.make_net(an)\n"
        "fr = Q.make_freq(an)\n"
        "fr._keep[2:] += %r\n"
        "D = Q.RowMatrix(np.zeros((an.nifo, an.N)))\n"
        "S = Q.RowMatrix(np.ones((an.nifo, an.N // 2)))\n"
        "p = W.injection0()\n"
        "rho = np.zeros(an.nifo)\n"
        "Q.lib().log_likelihood_full(ctypes.byref(net), D.p, Q._d(p), ctypes.byref(fr), S.p, Q._d(rho), an.N, an.Tobs)\n"
        % (ROOT, shift))
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = "-1"
    env["ROCR_VISIBLE_DEVICES"] = "-1"
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True, env=env, timeout=120)
    return r.returncode, r.stderr


def test_dropin_rejects_a_frequency_list_that_is_not_the_analysis_grid():
    """The drop-ins evaluate on f_i = i/Tobs (QuickCBC.c:417-419): a freq that
    differs is an invalid argument (the reference's error convention: print,
    exit(-1)), checked before any device work; the grid itself passes the
    check (and then, with no GPU, stops at the device)."""
    rc, err = _dropin_with_freq(0.5)
    assert rc == 255 and "log_likelihood_full: invalid argument" in err, (rc, err)
    rc, err = _dropin_with_freq(0.0)
    assert rc == 255 and "no usable gfx950 device" in err, (rc, err)
