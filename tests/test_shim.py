"""The SeqAn 2 drop-in header include/seqan_amd_seqan2.hpp (scores and, for
globalAlignment on StringSets of Gaps, the rows).

CPU: its overloads win overload resolution against the reference templates
(parallel_align_interface.h:332-365) for const, non-const and rvalue
arguments, and lose it for every call the library does not cover (static
assertions in tests/shim/shim_probe.cpp, compiled against the reference
headers); at run time a call reaches the library (which reports that no
device is present) and a length mismatch raises.
GPU: the same program, built beforehand, gives the reference CPU path's
scores value for value."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "tests", "shim")
RUNNER = os.path.join(SHIM, "_build", "shim_run")
REF_HEADERS = "/root/reference/include/seqan/align.h"
CASES = ["dna_global_nonconst", "dna_semiglobal_const", "dna_overlap_rvalue", "dna_local", "dna_concat_direct",
         "dna5_linear", "protein_local_blosum62", "protein_global_blosum62", "protein_semiglobal_blosum62",
         "protein_global_pam120", "protein_local_selectable_blosum80", "wavefront_global_semi",
         "wavefront_local_blosum62", "dna_rows", "protein_rows_blosum62", "dna_rows_semiglobal",
         "dna_rows_overlap", "dna_local_rows", "protein_local_rows_blosum62",
         "dna_rows_linear_gaps", "dna_rows_linear_gaps_semiglobal", "dna_local_rows_linear_gaps",
         "dna_banded", "dna_banded_semiglobal", "protein_banded_blosum62", "protein_local_banded_blosum62",
         "dna_banded_rows", "dna_banded_rows_semiglobal", "dna_local_banded_rows",
         "protein_local_banded_rows_blosum62", "dna_banded_one_diagonal", "dna_banded_corner_semiglobal",
         "dna_local_banded_corner", "dna_banded_rows_single_residue_h",
         "length_mismatch"]


def _run_cases():
    out = subprocess.run([RUNNER], capture_output=True, text=True, timeout=300, check=True).stdout
    print(out, flush=True)
    res = {}
    for line in out.splitlines():
        parts = line.split()
        if len(parts) >= 3 and parts[0] == "case":
            res[parts[1]] = parts[2]
    assert sorted(res) == sorted(CASES)
    return res


@pytest.mark.skipif(not os.path.exists(REF_HEADERS) or shutil.which("g++") is None,
                    reason="needs the reference headers and g++")
def test_overload_resolution_static():
    subprocess.run(["make", "-s", "-C", SHIM, "probe"], check=True, timeout=300)


@pytest.mark.skipif(not os.path.exists(RUNNER), reason="header test program not built (needs the reference headers)")
def test_overloads_reach_library_without_device():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present; covered by the GPU test")
    res = _run_cases()
    assert res.pop("length_mismatch") == "match"
    assert set(res.values()) == {"nodevice"}, res


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(RUNNER), reason="header test program not built (needs the reference headers)")
def test_overloads_match_reference_on_gpu():
    res = _run_cases()
    assert set(res.values()) == {"match"}, res
