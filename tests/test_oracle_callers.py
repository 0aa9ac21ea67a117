"""The reference-caller vectors (tests/golden/callers_golden.npz) reproduce
from the reference's own build (oracle/_ref/libgmapref_callers.so) where it
exists: the harness that tests/test_gpu_callers.py drives over the drop-in
gives, over the reference's SIMD fills, the committed results."""
import pytest

from test_gpu_callers import _compare, _compare_genome, _run

pytestmark = pytest.mark.ref


def test_reference_callers_reproduce_golden(reference, tmp_path):
    z, o = _run("ref", tmp_path)
    _compare(z, o)


def test_large_genome_reference_callers_reproduce_golden(reference, tmp_path):
    """The large-genome build of the reference gives the same results on a
    genome this small, so the standard build's vectors check both."""
    z, o = _run("ref_large", tmp_path)
    _compare(z, o)



def test_reference_genome_gap_reproduces_golden(reference, tmp_path):
    z, o = _run("ref", tmp_path, "genome")
    _compare_genome(z, o)


def test_large_genome_reference_genome_gap_reproduces_golden(reference, tmp_path):
    z, o = _run("ref_large", tmp_path, "genome")
    _compare_genome(z, o)


def test_reference_cdna_gap_reproduces_golden(reference, tmp_path):
    z, o = _run("ref", tmp_path, "cdna")
    _compare(z, o)


def test_large_genome_reference_cdna_gap_reproduces_golden(reference, tmp_path):
    z, o = _run("ref_large", tmp_path, "cdna")
    _compare(z, o)
