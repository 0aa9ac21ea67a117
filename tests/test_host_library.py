"""Host side of libgmapdp.so: exported C ABI, score tables, planning,
workload generator.  No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

import helpers as H
from gmapdp import lib
from gmapdp.lib import PROBLEM_DTYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gdp_[a-z0-9_]+|Dynprog_[A-Za-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def L():
    return lib.load()


def _elf_exports(path):
    """Defined global function symbols of an ELF64 shared object (.dynsym),
    read without loading it: libgmapdp_dynprog.so leaves the host program's
    tables and Pairpool_* undefined, so it only loads inside such a program."""
    import struct
    data = open(path, "rb").read()
    assert data[:4] == b"\x7fELF" and data[4] == 2
    shoff, = struct.unpack_from("<Q", data, 0x28)
    shentsize, shnum = struct.unpack_from("<HH", data, 0x3A)
    secs = [struct.unpack_from("<IIQQQQIIQQ", data, shoff + i * shentsize) for i in range(shnum)]
    out = set()
    for sh in secs:
        if sh[1] != 11:  # SHT_DYNSYM
            continue
        strtab = secs[sh[6]]
        for k in range(sh[5] // 24):
            name, info, _, shndx, _, _ = struct.unpack_from("<IBBHQQ", data, sh[4] + k * 24)
            if shndx != 0 and (info >> 4) in (1, 2) and (info & 15) == 2:  # defined GLOBAL/WEAK FUNC
                end = data.index(b"\0", strtab[4] + name)
                out.add(data[strtab[4] + name:end].decode())
    return out


def test_library_exports_every_declared_symbol(L):
    libs = {"gmapdp.h": lib.library_path(),
            "gmapdp_dynprog.h": os.path.join(os.path.dirname(lib.library_path()), "libgmapdp_dynprog.so")}
    headers = sorted(h for h in os.listdir(os.path.join(ROOT, "include")) if h.endswith(".h"))
    assert set(headers) == set(libs), headers
    for header in headers:
        names = _declared(header)
        assert names, header
        exported = _elf_exports(libs[header])
        for name in names:
            assert name in exported, "%s declares %s but %s does not export it" % (header, name, libs[header])
    assert {n for n in _elf_exports(libs["gmapdp.h"]) if n.startswith("gdp_")} >= set(_declared("gmapdp.h"))


def test_struct_sizes(L):
    assert L.gdp_descriptor_bytes() == 96
    assert PROBLEM_DTYPE.itemsize == ctypes.sizeof(lib.Problem) == 48
    assert ctypes.sizeof(lib.TbRequest) == 48 and ctypes.sizeof(lib.Pair) == 16


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5, 6])
def test_tables_match_oracle(L, mode):
    import pyoracle
    from gmapdp.lib import load
    O = pyoracle.Oracle(mode)
    L2 = load()
    L2.gdp_host_tables(mode)
    for mt in range(4):
        got = np.ctypeslib.as_array(L2.gdp_pairdistance(mt), shape=(128 * 128,)).reshape(128, 128)
        assert (got == O.pairdistance(mt)).all()
        assert L2.gdp_use8p_size(mt) == O.use8p(mt)
    assert (np.ctypeslib.as_array(L2.gdp_nt_to_int(), shape=(128,)) == O.nt_to_int()).all()
    for gs in range(3):
        for c in range(0, 128, 5):
            for g in range(128):
                assert L2.gdp_consistent(gs, c, g) == O.consistent(gs, c, g)


def test_compute_bands(L):
    lb, ub = ctypes.c_int32(), ctypes.c_int32()
    for rl, gl, eb, wide, want in [(100, 120, 10, 1, (10, 30)), (120, 100, 10, 1, (30, 10)), (5, 9, 3, 0, (3, 3))]:
        L.gdp_compute_bands(ctypes.byref(lb), ctypes.byref(ub), rl, gl, eb, wide)
        assert (lb.value, ub.value) == want


def test_plan_validates_and_sizes(L):
    from gmapdp.batch import plan
    specs = H.random_specs(9, 30)
    probs, q, g = H.pack_batch(specs, keep_scores=True)
    descs, dw, ss = plan(probs)
    assert dw > 0 and ss > 0
    bad = probs.copy()
    bad["width"][3] = 12
    with pytest.raises(RuntimeError, match="width"):
        plan(bad)
    # a band past the matrix is narrowed to it (no cell changes), so 700 +
    # 700 around a small problem plans; spans past GDP_MAX_SPAN after
    # narrowing do not
    wide = probs.copy()
    wide["lband"][0], wide["uband"][0], wide["kind"][0] = 700, 700, 0
    plan(wide)
    bad = probs.copy()
    bad["lband"][0], bad["uband"][0], bad["kind"][0] = 6000, 6000, 0
    bad["rl"][0], bad["gl"][0] = 5000, 5000
    with pytest.raises(RuntimeError, match="span"):
        plan(bad)
    bad = probs.copy()
    bad["open"][1] = 5
    with pytest.raises(RuntimeError, match="penalties"):
        plan(bad)
    bad = probs.copy()
    bad["kind"][4], bad["keep_scores"][4] = 0, 2  # the best end cell belongs to one-sided fills
    with pytest.raises(RuntimeError, match="upper or lower"):
        plan(bad)
    one = probs.copy()
    one["kind"], one["keep_scores"] = 1, 2
    assert plan(one)[2] == 4 * len(one)  # four score slots per best-cell record


def test_plan_validates_standard_fills(L):
    """Kind 3 (Dynprog_standard): width 32, its own flags and floor, a lower
    span cap; the SIMD kinds must leave those fields zero."""
    from gmapdp.batch import plan
    L.gdp_host_tables(0)
    specs = H.random_std_specs(12, 20)
    probs, _, _ = H.pack_batch(specs, keep_scores=True)
    descs, dw, ss = plan(probs)
    assert dw > 0 and ss > 0
    k = L.gdp_plan_kernels(descs.ctypes.data, len(probs))
    assert k & 0x10 and not k & 0xF  # the standard-fill launch only
    assert 6 <= (k >> 12) & 15 <= 10  # its ring: two words per column
    cases = [("width", 16, "width 32"), ("std_flags", 4, "std_flags"), ("saturation", -40000, "saturation")]
    for field, val, msg in cases:
        bad = probs.copy()
        bad[field][2] = val
        with pytest.raises(RuntimeError, match=msg):
            plan(bad)
    bad = probs.copy()
    bad["lband"][0], bad["uband"][0] = 264, 264
    bad["rl"][0], bad["gl"][0] = 400, 400  # a span past 526 after narrowing to the matrix
    with pytest.raises(RuntimeError, match="span"):
        plan(bad)
    simd, _, _ = H.pack_batch(H.random_specs(13, 4))
    simd["saturation"][1] = -5
    with pytest.raises(RuntimeError, match="GDP_KIND_STD"):
        plan(simd)


def test_cells_match_oracle(L, oracle):
    from gmapdp.batch import cells
    specs = H.random_specs(10, 60)
    probs, _, _ = H.pack_batch(specs)
    got = cells(probs)
    for k, s in enumerate(specs):
        want = {0: lambda: oracle.cells_full(s.rl, s.gl, s.lband, s.uband),
                1: lambda: oracle.cells_upper(s.rl, s.gl, s.uband),
                2: lambda: oracle.cells_lower(s.rl, s.gl, s.lband)}[s.kind]()
        assert got[k] == want


def test_workload_generator_is_deterministic_and_sliceable():
    from gmapdp.workload import Workload
    a = Workload(3, 400, genome_len=1 << 20)
    b = Workload(3, 100, genome_len=1 << 20, first=300)
    for k in range(100):
        pa, pb = a.problems[300 + k], b.problems[k]
        for f in ("rl", "gl", "lband", "uband", "width", "jump_late", "goff"):
            assert pa[f] == pb[f]
        qa = a.qbuf[pa["qoff"]: pa["qoff"] + pa["rl"]]
        qb = b.qbuf[pb["qoff"]: pb["qoff"] + pb["rl"]]
        assert (qa == qb).all()


def test_workload_parameter_ranges():
    from gmapdp.workload import Workload
    w = Workload(2, 3000, genome_len=1 << 20).problems
    assert (w["rl"] == 150).all() and w["gl"].min() >= 160 and w["gl"].max() <= 200
    assert w["lband"].min() >= 5 and w["lband"].max() <= 20 and (w["uband"] == w["gl"] - 150 + w["lband"]).all()
    assert (w["width"] == 16).all()
    w3 = Workload(3, 20000, genome_len=1 << 22).problems
    assert w3["rl"].min() >= 20 and w3["rl"].max() <= 1000
    r = w3["gl"] / w3["rl"]
    assert r.min() >= 0.99 and r.max() <= 1.51
    small = (w3["rl"] < 41) & (w3["gl"] < 41)
    assert ((w3["width"] == 8) == small).all()
    w4 = Workload(4, 2000, genome_len=1 << 20)
    p = w4.problems
    assert (p["kind"][0::2] == 1).all() and (p["kind"][1::2] == 2).all()
    assert ((p["width"] == 8) == ((p["rl"] < 24) | (p["gl"] < 24))).all()
    assert (p["mismatchtype"] == 3).all()


def test_plan_lane_group_parameters():
    from gmapdp.batch import plan
    specs = H.random_specs(3, 200, kinds=(0,), widths=(16,))
    probs, _, _ = H.pack_batch(specs)
    rl = probs["rl"]
    # the span the planner sweeps: bands narrowed to one past the matrix
    span = np.minimum(probs["lband"].astype(int), rl + 1) + np.minimum(probs["uband"].astype(int), probs["gl"] + 1)
    for params in (None, (240, 8001), (0, 8001), (64, 100), (240, 1)):
        descs, dw, ss = plan(probs, 0, params)
        d = descs[: len(probs) * 96].reshape(len(probs), 96)
        layout, glog = d[:, 58], d[:, 59]
        g16, g64 = (240, 8001) if params is None else params
        paired = span <= 960
        assert (layout[paired] == 1).all()
        want64 = paired & ((span > g16) | (rl >= g64))
        assert ((glog == 6) == want64)[paired].all(), params
        assert dw > 0
    with pytest.raises(RuntimeError):
        plan(probs, 0, (241, 8001))
    with pytest.raises(RuntimeError):
        plan(probs, 0, (240, 0))
