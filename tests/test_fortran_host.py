"""The reference program with its SODA branch on the GPU
(sodabayes_amd/fortran/build/sodabayes_gpu): same command line, .par files
and output files as the reference (SODABayes.f90).

The fixture tests/golden/fortran_jobs.npz holds whole reference jobs run
with -nthreads 1 (tests/golden/make_fortran_golden.py): inputs, arguments
and every output file.  With the default replay + FP64 path the GPU program
reproduces that chain, so the files agree to the last printed digit: integer
fields (iterations, model sizes, class counts) exactly, printed reals within
one unit of their last printed place (a last-bit difference of libm can move
a value across a rounding boundary).  Lines that carry wall-clock dates and
times are skipped.
"""
from __future__ import annotations

import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sodabayes_amd", "fortran", "build", "sodabayes_gpu")
FIXTURE = os.path.join(ROOT, "tests", "golden", "fortran_jobs.npz")
needs_exe = pytest.mark.skipif(not os.path.exists(EXE),
                               reason="the Fortran host is built only where the reference's Fortran sources are")


@pytest.fixture
def gpu_exe():
    """The GPU job tests are the drop-in boundary's end-to-end evidence: a
    missing program fails them instead of skipping them.  The program is
    built where the reference's Fortran sources are (build() of
    __graft_entry__) and travels with the tree."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built: run __graft_entry__.build() where the reference's Fortran "
                    "sources are, and ship the built program with the tree")
    return EXE
CLOCK_LINES = ("Run started at", "Run ended at", "Setup:")
SODAOFF_CASES = ("bayesr_sodaoff", "bayesa_sodaoff", "bayesb_sodaoff", "bayesrc_sodaoff", "bayescpi_sodaoff",
                 "bayesr5_sodaoff", "bayesr9_sodaoff", "bayesr_sodaoff_n1100", "bayescpi_sodaoff_n4500")
WIDE_MIX_CASES = ("bayesr5_blocks", "bayesr9_blocks",   # 5 and 9 classes, SODA blocks
                  "bayesr_blocks_n4500")                  # SODA blocks over three genotype chunks
# BayesA with -SODAOff: the .mod "Va" line prints varstore, which
# regular_bayes neither zeroes nor sums for BayesA (models.f90:650-655,
# 888-899), so the reference prints whatever its allocation held; the GPU
# host prints 0
UNDEFINED_LINES = {("bayesa_sodaoff", "ao.mod"): {3}}
NUM = re.compile(r"^[-+]?(\d+)?\.(\d*)(E([-+]\d+))?$")


def load_fixture():
    z = np.load(FIXTURE)
    manifest = json.loads(bytes(z["manifest"]).decode())
    return z, manifest


def last_place(tok: str) -> float | None:
    """One unit of the last printed digit of a real token, else None."""
    m = NUM.match(tok)
    if not m:
        return None
    return 10.0 ** ((int(m.group(4)) if m.group(4) else 0) - len(m.group(2)))


def compare_file(name: str, want: str, got: str, undefined=()) -> None:
    """Line by line; `undefined` lists (1-based) lines whose reference value
    is not defined by the reference's code, compared by their labels only."""
    wl = [ln for ln in want.splitlines() if not ln.lstrip().startswith(CLOCK_LINES)]
    gl = [ln for ln in got.splitlines() if not ln.lstrip().startswith(CLOCK_LINES)]
    assert len(wl) == len(gl), f"{name}: {len(gl)} lines, reference {len(wl)}"
    for k, (a, b) in enumerate(zip(wl, gl)):
        ta, tb = a.split(), b.split()
        if k + 1 in undefined:
            assert ta[:-1] == tb[:-1], f"{name}:{k + 1}: {b!r} vs reference {a!r}"
            continue
        assert len(ta) == len(tb), f"{name}:{k + 1}: {b!r} vs reference {a!r}"
        for x, y in zip(ta, tb):
            if x == y:
                continue
            u = last_place(x)
            assert u is not None and last_place(y) is not None, f"{name}:{k + 1}: {y!r} vs reference {x!r}"
            assert abs(float(x) - float(y)) <= 1.01 * u, f"{name}:{k + 1}: {y} vs reference {x}"


def write_inputs(z, case: str, files, d: str) -> None:
    for f in files:
        with open(os.path.join(d, f), "wb") as fh:
            fh.write(bytes(z[f"{case}/in/{f}"]))


def run(d: str, args, extra_env=None) -> subprocess.CompletedProcess:
    env = dict(os.environ, OMP_NUM_THREADS="1", **(extra_env or {}))
    r = subprocess.run([EXE, *args], cwd=d, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r


def test_fixture_lists_every_job():
    z, manifest = load_fixture()
    assert set(manifest) == {"bayesr_cmdline", "bayesa_rblocks", "bayesrc_groups", "bayescpi_nblocks",
                             "par_two_jobs", *SODAOFF_CASES, *WIDE_MIX_CASES}
    for case, m in manifest.items():
        for f in m["outputs"]:
            assert f"{case}/out/{f}" in z.files
        assert any(f.endswith(".hyp") for f in m["outputs"])
    # the mixture job's trace carries the class counts N1..N4
    hyp = bytes(z["bayesr_cmdline/out/r.hyp"]).decode().splitlines()
    assert hyp[0].split()[:2] == ["Iteration", "ModelSize"] and len(hyp) == 31


def test_last_place_reads_the_printed_precision():
    assert last_place("0.100056E+03") == pytest.approx(1e-3)
    assert last_place("0.5000000E-02") == pytest.approx(1e-9)
    assert last_place("0.123456") == pytest.approx(1e-6)
    assert last_place("181") is None and last_place("snp12") is None
    compare_file("x", "a 0.100056E+03 7\n", "a 0.100057E+03 7\n")
    compare_file("x", "Va 0.477047-314\n", "Va 0.000000E+00\n", undefined={1})
    with pytest.raises(AssertionError):
        compare_file("x", "a 0.100056E+03 7\n", "a 0.100058E+03 7\n")
    with pytest.raises(AssertionError):
        compare_file("x", "a 0.100056E+03 7\n", "a 0.100056E+03 8\n")


@needs_exe
def test_help_lists_the_reference_options():
    """No arguments: the reference's option table (useroptions.f90:24-54,
    printed by parse_help), without touching the GPU."""
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    keys = [ln.split()[0] for ln in r.stdout.splitlines()[2:] if ln.strip()]
    assert keys[:3] == ["-parfile", "-job_name", "-method"] and keys[-1] == "-out" and len(keys) == 31


@pytest.mark.gpu
@pytest.mark.usefixtures("gpu_exe")
@pytest.mark.parametrize("case,dense_x", [("bayesr_cmdline", False), ("bayesa_rblocks", False),
                                          ("bayesrc_groups", False), ("bayescpi_nblocks", False),
                                          ("par_two_jobs", False), ("bayesr_cmdline", True),
                                          *((c, False) for c in WIDE_MIX_CASES)])
def test_gpu_job_files_match_the_reference(tmp_path, case, dense_x):
    """By default the program reads the .bed packed and computes the allele
    frequencies from its counts (no dense genotype matrix on the host);
    SODABAYES_HOST_DENSE_X keeps the reference's loaders.  Both give the
    reference's files."""
    z, manifest = load_fixture()
    m = manifest[case]
    write_inputs(z, case, m["inputs"], str(tmp_path))
    r = run(str(tmp_path), m["args"], {"SODABAYES_HOST_DENSE_X": "1"} if dense_x else None)
    assert "SODA on the GPU: rng replay , precision fp64 , chains 1" in r.stdout
    for f in m["outputs"]:
        path = tmp_path / f
        assert path.exists(), f"{f} not written"
        compare_file(f, bytes(z[f"{case}/out/{f}"]).decode(), path.read_text())


@pytest.mark.gpu
@pytest.mark.usefixtures("gpu_exe")
@pytest.mark.parametrize("case", SODAOFF_CASES)
def test_gpu_single_site_job_files_match_the_reference(tmp_path, case):
    """-SODAOff: the single-site sampler (models.f90:607-962) on the GPU, one
    workgroup walking the SNPs in id order with the reference's stream-0
    draws, gives the reference's files."""
    z, manifest = load_fixture()
    m = manifest[case]
    write_inputs(z, case, m["inputs"], str(tmp_path))
    r = run(str(tmp_path), m["args"])
    assert "single-site sampler on the GPU: rng replay , chains 1" in r.stdout
    for f in m["outputs"]:
        undefined = UNDEFINED_LINES.get((case, f), ())
        compare_file(f, bytes(z[f"{case}/out/{f}"]).decode(), (tmp_path / f).read_text(), undefined)
    if case == "bayesa_sodaoff":
        assert (tmp_path / "ao.mod").read_text().splitlines()[2].split() == ["Va", "0.000000E+00"]


@needs_exe
def test_single_site_job_on_the_host_matches_the_reference(tmp_path):
    """SODABAYES_REGULAR_HOST keeps the reference's host sampler and dense
    loaders for -SODAOff; no device call is made."""
    z, manifest = load_fixture()
    m = manifest["bayesr_sodaoff"]
    write_inputs(z, "bayesr_sodaoff", m["inputs"], str(tmp_path))
    r = run(str(tmp_path), m["args"], {"SODABAYES_REGULAR_HOST": "1"})
    assert "on the GPU" not in r.stdout
    for f in m["outputs"]:
        compare_file(f, bytes(z[f"bayesr_sodaoff/out/{f}"]).decode(), (tmp_path / f).read_text())


@pytest.mark.gpu
@pytest.mark.usefixtures("gpu_exe")
def test_gpu_default_partition_of_one_snp_runs_the_single_site_sampler(tmp_path):
    """Neither -blocksize nor -nblocks with -nthreads 1 partitions into
    blocks of one SNP: setup_soda switches SODA off and the reference then
    crashes in soda_bayes (soda.f90:418-421, SODABayes.f90:134-142).  Here
    the job runs the single-site sampler, and its files equal the
    reference's -SODAOff files for the same job."""
    z, manifest = load_fixture()
    m = manifest["bayescpi_sodaoff"]
    write_inputs(z, "bayescpi_sodaoff", m["inputs"], str(tmp_path))
    args = [a for a in m["args"] if a != "-SODAOff"]
    r = run(str(tmp_path), args)
    assert "single-site sampler on the GPU" in r.stdout
    for f in m["outputs"]:
        compare_file(f, bytes(z[f"bayescpi_sodaoff/out/{f}"]).decode(), (tmp_path / f).read_text())


@pytest.mark.gpu
@pytest.mark.usefixtures("gpu_exe")
def test_gpu_single_site_counter_chains(tmp_path):
    """-SODAOff with SODABAYES_RNG=counter and 3 chains: one set of files per
    chain, independent chains, each near the reference's posterior mean of
    mu."""
    z, manifest = load_fixture()
    m = manifest["bayesr_sodaoff"]
    write_inputs(z, "bayesr_sodaoff", m["inputs"], str(tmp_path))
    r = run(str(tmp_path), m["args"], {"SODABAYES_RNG": "counter", "SODABAYES_CHAINS": "3"})
    assert "single-site sampler on the GPU: rng counter, chains 3" in r.stdout
    ref_mu = float(bytes(z["bayesr_sodaoff/out/ro.mod"]).decode().split()[1])
    effs = []
    for sfx in ("", "_chain1", "_chain2"):
        lines = (tmp_path / f"ro{sfx}.eff").read_text().splitlines()
        effs.append(np.array([float(ln.split()[1]) for ln in lines[1:]]))
        mod = (tmp_path / f"ro{sfx}.mod").read_text().split()
        assert mod[0] == "Mean" and abs(float(mod[1]) - ref_mu) < 0.01 * abs(ref_mu)
    assert not np.array_equal(effs[0], effs[1]) and not np.array_equal(effs[1], effs[2])


@pytest.mark.gpu
@pytest.mark.usefixtures("gpu_exe")
def test_gpu_predict_job_matches_the_reference(tmp_path):
    """-predict on the reference's model files: genomic values through
    sb_predict in place of compute_dgv."""
    z, manifest = load_fixture()
    m = manifest["bayesr_cmdline"]
    write_inputs(z, "bayesr_cmdline", m["inputs"], str(tmp_path))
    for f in ("r.mod", "r.frq", "r.eff"):
        (tmp_path / f).write_bytes(bytes(z[f"bayesr_cmdline/out/{f}"]))
    run(str(tmp_path), m["predict_args"])
    for f in m["predict_outputs"]:
        compare_file(f, bytes(z[f"bayesr_cmdline/out/predict/{f}"]).decode(), (tmp_path / f).read_text())


@pytest.mark.gpu
@pytest.mark.usefixtures("gpu_exe")
def test_gpu_batched_chains_write_one_set_of_files_each(tmp_path):
    """SODABAYES_CHAINS=4 with the counter-based mixed path: the usual files
    for chain 0, <out>_chain<k>.eff/.mod/.gv for the others, each a complete
    posterior summary of its own chain."""
    z, manifest = load_fixture()
    m = manifest["bayesr_cmdline"]
    write_inputs(z, "bayesr_cmdline", m["inputs"], str(tmp_path))
    r = run(str(tmp_path), m["args"], {"SODABAYES_RNG": "counter", "SODABAYES_PRECISION": "mixed",
                                       "SODABAYES_CHAINS": "4"})
    assert "rng counter, precision mixed, chains 4" in r.stdout
    ref_eff = bytes(z["bayesr_cmdline/out/r.eff"]).decode().splitlines()
    effs = []
    for sfx in ("", "_chain1", "_chain2", "_chain3"):
        for ext in (".eff", ".mod", ".gv"):
            assert (tmp_path / f"r{sfx}{ext}").exists(), f"r{sfx}{ext}"
        lines = (tmp_path / f"r{sfx}.eff").read_text().splitlines()
        assert lines[0].split() == ref_eff[0].split() and len(lines) == len(ref_eff)
        effs.append(np.array([float(ln.split()[1]) for ln in lines[1:]]))
        pip = np.array([[float(v) for v in ln.split()[2:]] for ln in lines[1:]])
        assert np.allclose(pip.sum(1), 1.0, atol=1e-5)   # class probabilities per SNP
    assert not np.array_equal(effs[0], effs[1])           # independent chains
    # posterior mean of mu (.mod "Mean") of every chain near the reference's
    ref_mu = float(bytes(z["bayesr_cmdline/out/r.mod"]).decode().split()[1])
    for sfx in ("", "_chain1", "_chain2", "_chain3"):
        mod = (tmp_path / f"r{sfx}.mod").read_text().split()
        assert mod[0] == "Mean" and abs(float(mod[1]) - ref_mu) < 0.01 * abs(ref_mu)
