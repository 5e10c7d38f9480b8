"""CPU checks of the C-ABI library: it loads, exports every function declared
in include/*.h, its structs match the header byte for byte, and argument
validation fails cleanly without touching the GPU."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dida_letkf.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src)
    skip = {"if", "sizeof", "defined"}
    # declarations end with ');' -- keep names that are followed by a prototype
    out = []
    for m in re.finditer(r"(?:^|\n)\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ \*]*?\b([A-Za-z_]\w*)\s*\(([^;{]*?)\)\s*;",
                         src):
        if m.group(1) not in skip:
            out.append(m.group(1))
    return sorted(set(out))


@pytest.fixture(scope="module")
def L():
    from dida_letkf import engine
    return engine.lib()


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for n in ("letkf_core_", "dida_letkf_create", "dida_letkf_analyze", "dida_letkf_set_obs",
              "dida_letkf_core_batched", "dida_letkf_epilogue_device"):
        assert n in names


def test_library_exports_every_declared_symbol(L):
    for n in declared_functions():
        assert hasattr(L, n), f"missing export {n}"


def test_python_binding_lists_every_export():
    from dida_letkf import engine
    assert sorted(engine.EXPORTS) == declared_functions()


def _c_layout():
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "dida_letkf.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(dida_letkf_config), offsetof(dida_letkf_config, local_dist),
         offsetof(dida_letkf_config, hyai), offsetof(dida_letkf_config, keep_transforms),
         sizeof(dida_letkf_stats), offsetof(dida_letkf_stats, ms_select),
         sizeof(dida_serve_job), offsetof(dida_serve_job, obs), sizeof(dida_serve_obs),
         sizeof(dida_serve_status));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(prog)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        return [int(x) for x in subprocess.run([exe], capture_output=True, text=True).stdout.split()]


def test_struct_layout_matches_header():
    from dida_letkf.engine import DidaConfig, DidaStats
    s = _c_layout()
    assert s[0] == C.sizeof(DidaConfig)
    assert s[1] == DidaConfig.local_dist.offset
    assert s[2] == DidaConfig.hyai.offset
    assert s[3] == DidaConfig.keep_transforms.offset
    assert s[4] == C.sizeof(DidaStats)
    assert s[5] == DidaStats.ms_select.offset
    from dida_letkf.engine import ServeJob, ServeObs, ServeStatus
    assert s[6] == C.sizeof(ServeJob) and s[7] == ServeJob.obs.offset
    assert s[8] == C.sizeof(ServeObs) and s[9] == C.sizeof(ServeStatus)


def test_n_analyses_matches_config(L):
    from dida_letkf import synth
    from dida_letkf.engine import to_c_config
    for name, kw in (("c0", {}), ("c1", dict(nlat=18, nlon=36, nz=4)), ("c2", {})):
        cfg = synth.make_config(name, **kw)
        c = to_c_config(cfg)
        assert L.dida_letkf_n_analyses(C.byref(c)) == cfg.n_analyses()


def test_invalid_config_is_rejected_before_any_gpu_call(L):
    from dida_letkf import synth
    from dida_letkf.engine import to_c_config, DidaError, _check
    cfg = synth.make_config("c0")
    for field, bad in (("k", 1), ("k", 500), ("local_dist", 0.0), ("inflation_factor", -1.0),
                       ("precision", 7), ("nlat", 0)):
        c = to_c_config(cfg)
        setattr(c, field, bad)
        h = C.c_void_p()
        rc = L.dida_letkf_create(C.byref(c), C.byref(h))
        assert rc == -1 or rc == -6, (field, rc)
        assert not h.value
        assert L.dida_letkf_last_error()
    with pytest.raises(DidaError):
        _check(-1, "x")


def test_core_batched_rejects_bad_rows(L):
    import numpy as np
    rp = np.array([0, 5, 3], np.int64)
    tr = np.zeros(2 * 4)
    w = np.zeros(2 * 2)
    rc = L.dida_letkf_core_batched(2, 2, rp.ctypes.data_as(C.POINTER(C.c_longlong)), None, None,
                                   None, None, tr.ctypes.data_as(C.POINTER(C.c_double)),
                                   w.ctypes.data_as(C.POINTER(C.c_double)), 1.0, 0, None)
    assert rc == -1
