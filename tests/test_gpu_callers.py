"""The drop-in driven by the reference's own callers.

oracle/_ref/libgmapref_callers_gpu.so is the reference's dynprog_single.c,
dynprog_end.c, dynprog_cdna.c and dynprog_genome.c (with dynprog.c, pairpool.c, genome.c,
maxent_hr.c, pair.c), compiled in place and linked against
libgmapdp_dynprog.so instead of dynprog_simd.o.  Its Dynprog_single_gap,
Dynprog_end5_gap and Dynprog_end3_gap, called with the arguments of
tests/golden/callers_golden.npz, and its Dynprog_genome_gap and
Dynprog_cdna_gap, called with those of tests/golden/callers_genome_golden.npz
and callers_cdna_golden.npz, must return exactly the pair lists and counters the same callers returned over the reference's own
SIMD fills (make_callers_golden.py): the fills, end-cell scans of the returned
tables, walks, genome skips through the callers' Genome_T, and the callers'
own length chopping, shortcuts and post-filters all run as in GMAP.  The
calls run in a process of their own (tests/callers_run.py), where the
drop-in resolves its host symbols against the callers' build alone.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.ref]
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "callers_golden.npz")


def _compare(z, o):
    npairs = z["out_npairs"]
    assert (o["npairs"] == npairs).all(), np.nonzero(o["npairs"] != npairs)[0][:10]
    assert (o["counters"] == z["out_counters"]).all(), np.nonzero((o["counters"] != z["out_counters"]).any(1))[0][:10]
    assert o["pairs"].shape == z["out_pairs"].shape and (o["pairs"] == z["out_pairs"]).all()


GOLD_EXTRA = {"genome": os.path.join(HERE, "golden", "callers_genome_golden.npz"),
              "cdna": os.path.join(HERE, "golden", "callers_cdna_golden.npz")}


def _run(which, tmp_path, kind=""):
    from callers_run import callers_lib
    lib = callers_lib(which)
    if not os.path.exists(lib):
        pytest.skip("%s not built" % os.path.basename(lib))
    z = np.load(GOLD_EXTRA.get(kind, GOLD), allow_pickle=False)
    cases = tmp_path / "cases.npz"
    np.savez(cases, **{k: z[k] for k in z.files if not k.startswith("out_")})
    out = tmp_path / "out.npz"
    print("running the reference's %s callers over the %s fills" % (kind or "single/end", which), flush=True)
    r = subprocess.run([sys.executable, os.path.join(HERE, "callers_run.py"), which, str(cases), str(out)] +
                       ([kind] if kind else []), capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    return z, np.load(out)


def _compare_genome(z, o):
    _compare(z, o)
    assert (o["probs"] == z["out_probs"]).all(), np.nonzero((o["probs"] != z["out_probs"]).any(1))[0][:10]


def test_reference_callers_over_the_dropin(tmp_path):
    z, o = _run("gpu", tmp_path)
    _compare(z, o)
    # the vectors cover both callers, both widths and NULL lists
    assert z["single"].sum() >= 1500 and (~z["single"].astype(bool)).sum() >= 800


def test_reference_callers_over_the_large_genome_dropin(tmp_path):
    """GMAP's large-genome configuration (gmapl, LARGE_GENOMES: 64-bit
    Univcoord_T in the tracebacks' chromosome bounds): the callers built so,
    over libgmapdp_dynprog_large.so, return what the standard build's
    callers over the reference's fills returned."""
    z, o = _run("gpu_large", tmp_path)
    _compare(z, o)


def test_reference_genome_gap_over_the_dropin(tmp_path):
    """Dynprog_genome_gap (dynprog_genome.c:3287-3770) over the drop-in: its
    four one-sided fills and their score tables, the reference's own intron
    bridge over those tables, both ends' walks and the gap holder between
    them, as over the reference's fills."""
    z, o = _run("gpu", tmp_path, "genome")
    _compare_genome(z, o)
    # the vectors hold joined introns, NULL lists and both widths' fills
    assert (z["out_npairs"] > 0).sum() >= 400 and (z["out_npairs"] < 0).sum() >= 40


def test_reference_genome_gap_over_the_large_genome_dropin(tmp_path):
    z, o = _run("gpu_large", tmp_path, "genome")
    _compare_genome(z, o)


def test_reference_cdna_gap_over_the_dropin(tmp_path):
    """Dynprog_cdna_gap (dynprog_cdna.c:786-1180) over the drop-in: the four
    one-sided fills and their score tables, the reference's own cDNA bridge
    over them, both ends' walks and the gap holder left for the insertion."""
    z, o = _run("gpu", tmp_path, "cdna")
    _compare(z, o)
    assert (z["out_npairs"] > 0).sum() >= 500 and (z["out_npairs"] < 0).sum() >= 5


def test_reference_cdna_gap_over_the_large_genome_dropin(tmp_path):
    z, o = _run("gpu_large", tmp_path, "cdna")
    _compare(z, o)
