#!/usr/bin/env python3
"""Writes tests/golden/callers_golden.npz: calls of the reference's own
callers Dynprog_single_gap (dynprog_single.c:428-675), Dynprog_end5_gap
(dynprog_end.c:1293-1650) and Dynprog_end3_gap (1923-2243) over the
reference's SIMD fills, with the genome-skip test genome (the reference's
test chromosome with N runs) cut into two chromosomes; and
tests/golden/callers_genome_golden.npz: 600 calls of Dynprog_genome_gap
(dynprog_genome.c:3287-3770) from helpers.genome_gap_calls (seed 1002) over
the same genome, with their pair lists (the gap holder carries the intron),
counters {dynprogindex, new_leftgenomepos, new_rightgenomepos,
traceback_score, nmatches, nmismatches, nopens, nindels, exonhead,
introntype} and splice-site probabilities {left_prob, right_prob}; and
tests/golden/callers_cdna_golden.npz: 600 calls of Dynprog_cdna_gap
(dynprog_cdna.c:786-1180) from helpers.cdna_gap_calls (seed 1003), with
their pair lists (the gap holder between the ends marks the cDNA insertion)
and counters {dynprogindex, traceback_score, incompletep}; and
tests/golden/callers_cdna_insert_golden.npz: the same for calls whose join
takes the insertion-pairs branch (main_cdna_insert).  With the argument
`genome`, `cdna` or `cdna_insert` only that file is written.

Run in a container where the reference tree is mounted, after
`make -C oracle ref`, which builds oracle/_ref/libgmapref_callers.so from
the reference sources in place.  The cases are helpers.caller_cases: 2000
single gaps of the single-gap generator (HIGHQ int16 and LOWQ int8 fills,
plus about 2000-column bands and ungapped windows) and 1000 end gaps of the
end-extension generator under all four end modes.  Stored: the calls'
arguments and reads, and the returned pair lists (8 ints per pair, from the
list's head) and counters {dynprogindex, traceback_score, nmatches,
nmismatches, nopens, nindels}; npairs -1 is a NULL list.
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "gmap-gsnap_amd")]

import callers_run  # noqa: E402
import golden_io  # noqa: E402
import helpers as H  # noqa: E402

OUT = os.path.join(HERE, "callers_golden.npz")
OUT_GENOME = os.path.join(HERE, "callers_genome_golden.npz")
OUT_CDNA = os.path.join(HERE, "callers_cdna_golden.npz")
OUT_CDNA_INSERT = os.path.join(HERE, "callers_cdna_insert_golden.npz")
SEED, SEED_GENOME, SEED_CDNA, N_EXTRA = 1001, 1002, 1003, 600
SEED_CDNA_INSERT, N_INSERT = 1004, 300


def main():
    seq, _ = golden_io.skip_genome_image()
    if sys.argv[1:] in (["genome"], ["cdna"]):
        return main_extra(seq, sys.argv[1])
    if sys.argv[1:] == ["cdna_insert"]:
        return main_cdna_insert(seq)
    cases = H.caller_cases(np.random.default_rng(SEED), seq, 2000, 1000)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath = os.path.join(d, "cases.npz"), os.path.join(d, "out.npz")
        callers_run.save_cases(cpath, cases)
        callers_run.run("ref", cpath, opath)
        z, o = np.load(cpath), np.load(opath)
        np.savez_compressed(OUT, **{k: z[k] for k in z.files}, **{"out_" + k: o[k] for k in o.files})
    print("wrote %s: %d calls" % (OUT, len(cases)))
    main_extra(seq, "genome")
    main_extra(seq, "cdna")
    main_cdna_insert(seq)


def main_extra(seq, kind):
    if kind == "genome":
        cases = H.genome_gap_calls(np.random.default_rng(SEED_GENOME), seq, N_EXTRA)
        save, run, out = callers_run.save_genome_cases, callers_run.run_genome, OUT_GENOME
    else:
        cases = H.cdna_gap_calls(np.random.default_rng(SEED_CDNA), seq, N_EXTRA)
        save, run, out = callers_run.save_cdna_cases, callers_run.run_cdna, OUT_CDNA
    with tempfile.TemporaryDirectory() as d:
        cpath, opath = os.path.join(d, "cases.npz"), os.path.join(d, "out.npz")
        save(cpath, cases)
        run("ref", cpath, opath)
        z, o = np.load(cpath), np.load(opath)
        np.savez_compressed(out, **{k: z[k] for k in z.files}, **{"out_" + k: o[k] for k in o.files})
    print("wrote %s: %d calls" % (out, len(cases)))


def main_cdna_insert(seq):
    """Dynprog_cdna_gap calls whose join leaves 9 query and 9 genome positions
    (INSERT_PAIRS, dynprog_cdna.c:1007-1030): query gaps as long as their
    genome gaps with 9 bases replaced (helpers.cdna_gap_calls, block 9).  Run
    one after another in one process, some of these calls leave the
    reference's caller in a state that crashes a later call, while every call
    finishes in a fresh process; so each call runs in a child process of its
    own and the outputs are joined in order."""
    import subprocess
    cases = H.cdna_gap_calls(np.random.default_rng(SEED_CDNA_INSERT), seq, N_INSERT, block=9)
    outs = []
    with tempfile.TemporaryDirectory() as d:
        for k in range(len(cases)):
            cpath, opath = os.path.join(d, "c.npz"), os.path.join(d, "o.npz")
            callers_run.save_cdna_cases(cpath, cases[k:k + 1])
            code = ("import sys; sys.path[:0] = %r; import callers_run; callers_run.run_cdna('ref', %r, %r)"
                    % (sys.path[:3], cpath, opath))
            subprocess.run([sys.executable, "-c", code], stdout=subprocess.DEVNULL, check=True)
            o = np.load(opath)
            outs.append({f: o[f] for f in o.files})
        cpath = os.path.join(d, "cases.npz")
        callers_run.save_cdna_cases(cpath, cases)
        z = np.load(cpath)
        out = {"out_" + f: np.concatenate([x[f] for x in outs]) for f in outs[0]}
        np.savez_compressed(OUT_CDNA_INSERT, **{k: z[k] for k in z.files}, **out)
    print("wrote %s: %d calls" % (OUT_CDNA_INSERT, len(cases)))


if __name__ == "__main__":
    main()
