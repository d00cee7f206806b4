"""The C ABI library: it loads without a GPU, exports exactly the entry points that
include/pyjuice_amd.h declares, the ctypes table matches the header, and argument checks
reject bad sizes on the host before any launch."""
import ctypes
import os
import re

import pytest

from pyjuice_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pyjuice_amd.h")

_CTYPE = {"int": N._i32, "int64_t": N._i64, "float": N._f32}


def _declarations():
    """name -> list of parameter types (as written) for every function in the header."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = dict()
    for m in re.finditer(r"^(?:const\s+)?\w+\*?\s+(pjamd_\w+)\(([^)]*)\);", text, flags=re.M):
        params = [p.strip() for p in m.group(2).replace("\n", " ").split(",") if p.strip() and p.strip() != "void"]
        decls[m.group(1)] = params
    return decls


def _ctype_of(param: str):
    t = param.rsplit(" ", 1)[0].strip() if not param.endswith("*") else param
    if "*" in param or "pjamd_stream_t" in param:
        return N._c_p
    return _CTYPE[t.replace("const ", "").strip()]


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(N.lib_path()):
        pytest.skip("libpyjuice_amd.so is not built")
    return N.lib()


def test_header_parses():
    d = _declarations()
    assert len(d) >= 20
    assert "pjamd_sum_forward" in d and "pjamd_sum_backward_par_batched" in d


def test_library_exports_every_declared_symbol(lib):
    for name in _declarations():
        assert hasattr(lib, name), f"{name} is declared in the header but not exported"


def test_library_exports_nothing_else(lib):
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", N.lib_path()], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("nm unavailable")
    exported = {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}
    assert exported == set(_declarations()), exported ^ set(_declarations())


def test_ctypes_table_matches_header():
    decls = _declarations()
    for name, params in decls.items():
        if name == "pjamd_error_string":
            continue
        assert name in N._SIGNATURES, f"{name} has no ctypes signature"
        want = [_ctype_of(p) for p in params]
        assert N._SIGNATURES[name] == want, (name, N._SIGNATURES[name], want)


def test_table_widths_match_header():
    import re
    from pyjuice_amd.model.tensorcircuit import PAR_DESC_FIELDS
    text = open(HEADER).read()
    assert int(re.search(r"#define PJAMD_PAR_DESC_FIELDS (\d+)", text).group(1)) == PAR_DESC_FIELDS


def test_host_only_entry_points(lib):
    assert lib.pjamd_abi_version() == 1
    assert lib.pjamd_error_string(0) is not None
    assert lib.pjamd_error_string(-1) is not None


def test_argument_checks_fail_before_launch(lib):
    null = ctypes.c_void_p(0)
    # negative block count / zero batch: rejected on the host with PJAMD_STATUS_EINVAL
    assert lib.pjamd_prod_forward(null, null, null, null, -1, 2, 4, 8, 0, null) == -1
    assert lib.pjamd_sum_forward(null, null, null, null, null, null, null, 1, 4, 32, 0, 0, 1, null) == -1
    assert lib.pjamd_sum_forward(null, null, null, null, null, null, null, 1, 4, 32, 8, 0, 7, null) == -1
    assert lib.pjamd_sum_backward_par_batched(null, null, null, null, null, null, null, null, null, null, null,
                                              null, 4, 48, 8, 0, 1, null) == -1
    # empty work is a no-op that succeeds without touching the device
    assert lib.pjamd_prod_forward(null, null, null, null, 0, 2, 4, 8, 0, null) == 0
    assert lib.pjamd_sum_backward_par_batched(null, null, null, null, null, null, null, null, null, null, null,
                                              null, 0, 64, 8, 0, 1, null) == 0


def test_argument_checks_of_query_and_split_entry_points(lib):
    null = ctypes.c_void_p(0)
    p7, p9 = [null] * 7, [null] * 9
    # propagation algorithms: unknown alg, GeneralLL without a positive alpha, zero batch
    assert lib.pjamd_sum_forward_alg(*p7, 1, 4, 32, 8, 3, 1.0, null) == -1
    assert lib.pjamd_sum_forward_alg(*p7, 1, 4, 32, 8, 2, 0.0, null) == -1
    assert lib.pjamd_sum_forward_alg(*p7, 0, 4, 32, 8, 1, 0.0, null) == 0
    assert lib.pjamd_sum_backward_ele_alg(*p9, 1, 2, 32, 32, 8, 2, -1.0, 0, null) == -1
    assert lib.pjamd_sum_backward_par_alg(null, 0, *p9, 1, 4, 32, 0, 1, null) == -1
    # tied-range input backward: Categorical only, category counts up to 16384
    assert lib.pjamd_input_backward_tied(1, null, null, 0, *([null] * 6), 4, 2, 0, 8, 0, null) == -2
    assert lib.pjamd_input_backward_tied(0, null, null, 0, *([null] * 6), 4, 20000, 0, 8, 0, null) == -2
    assert lib.pjamd_input_backward_tied(0, null, null, 0, *([null] * 6), 0, 16, 0, 8, 0, null) == 0
    # split reductions: 2 .. 16 splits
    assert lib.pjamd_sum_forward_split(*p7, 1, 1024, 1024, 256, 1024, 1, null, 1, null) == -1
    assert lib.pjamd_sum_forward_split(*p7, 1, 1024, 1024, 256, 1024, 17, null, 1, null) == -1
