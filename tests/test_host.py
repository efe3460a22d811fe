"""Host logic and the C ABI library on the CPU (no kernel launches)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import orc
from common import load_case, soda_cases
from sodabayes_amd import _lib, soda

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "sodabayes.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sb_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert set(names) == set(_lib.EXPORTS)
    for f in names:
        assert hasattr(L, f), f


@pytest.mark.parametrize("case", soda_cases())
def test_partition_matches_reference(case):
    c = load_case(case)
    o = c["opts"]
    bs, nb, bt = soda.partition(int(c["nloci"]), o.get("blocksize", 0), o.get("nblocks", 0), 1)
    assert (bs, nb, bt) == (int(c["out"]["blocksize"][0]), int(c["out"]["nblocks"][0]),
                            int(c["out"]["blockt"][0]))
    bounds = soda.block_bounds(int(c["nloci"]), bs, nb, bt)
    assert bounds[0][0] == 0 and bounds[-1][1] == int(c["nloci"])
    assert all(a[1] == b[0] for a, b in zip(bounds, bounds[1:]))


def test_host_rng_seed_and_permutation_match_reference():
    for seed in (11, -42, 123456789):
        st = soda.rng_seed(seed)
        r = orc.Rng(seed)
        assert tuple(int(v) for v in st) == r.state()
        p = soda.permutate(777, 64, st)
        q = orc.permutate(777, 64, r)
        assert np.array_equal(p + 1, q)
        assert tuple(int(v) for v in st) == r.state()


def test_rblocks_permutation_matches_golden():
    c = load_case("rblocks")
    o = c["opts"]
    st = soda.rng_seed(o["seed"])
    p = soda.permutate(int(c["nloci"]), o["blocksize"], st)
    assert np.array_equal(p + 1, c["out"]["permvec"])


def test_options_follow_reference_parsing():
    o = soda.SodaOptions(method="bayesr", nmix=4, mix="0.0,0.0001,0.001,0.01", scale_va=0.01)
    assert o.method_code() == 5
    assert o.mix_values()[1] == float(np.float32(0.0001))
    assert soda.f32(0.01) != 0.01
    assert soda.SodaOptions(method="BayesCpi", pi=0.5).mix_values() == [0.0, 1.0]
    assert soda.SodaOptions(method="BayesA").resolved_nmix() == 1
    with pytest.raises(ValueError):
        soda.SodaOptions(method="BayesX").method_code()
    with pytest.raises(ValueError):
        soda.SodaOptions(method="BayesR", nmix=3, mix="0,1").mix_values()


def test_group_index_by_first_appearance():
    idx, ng = soda.group_index(["3", "1", "3", "2", "1"])
    assert ng == 3 and list(idx) == [0, 1, 0, 2, 1]


def test_library_errors_are_reported():
    L = _lib.load()
    with pytest.raises(_lib.SbError):
        _lib.check(L.sb_rng_seed(1, -1, None))
    assert b"sb_rng_seed" in L.sb_last_error()


def test_fortran_binding_calls_the_library():
    """The iso_c_binding module (sodabayes_amd/fortran/soda_gpu.f90) compiled
    with the image's Fortran compiler calls the library's host entry points."""
    import subprocess
    from sodabayes_amd import build as b
    exe = os.path.join(b.FBUILD, "binding_check")
    if not os.path.exists(exe):
        exe = b.build_fortran()
        if exe is None:
            pytest.skip("no Fortran compiler")
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    r = orc.Rng(123456789)
    s0 = [int(v) if v < 2**63 else int(v) - 2**64 for v in r.state()]
    assert [int(t) for t in out[0].split()] == s0
    perm = orc.permutate(1000, 64, r) - 1
    assert [int(t) for t in out[1].split()] == list(perm[:10])
    s1 = [int(v) if v < 2**63 else int(v) - 2**64 for v in r.state()]
    assert [int(t) for t in out[2].split()] == s1
    assert out[3].startswith("-1 sb_rng_seed")
