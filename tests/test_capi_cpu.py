"""The C ABI library loads, exports every declared entry point and rejects bad
input before touching a device.  CPU only: no compute calls."""
import ctypes as C
import os
import re
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bf", "bf.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ["bf_create", "bf_destroy", "bf_init_fields", "bf_step", "bf_read_field", "bf_conservation"]:
        assert s in syms


def test_library_exports_every_declared_symbol(bfa):
    lib = bfa.load_library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_library_loaded_from_a_path_serves_the_process(bfa):
    """A library loaded from an explicit path (a variant build) is the one the
    Solver class and the module functions call afterwards."""
    from boltzmannfluids_amd import capi
    default = bfa.load_library()
    try:
        other = bfa.load_library(capi.LIB_PATH)
        assert other is not default
        assert bfa.load_library() is other
        assert bfa.velocity_table("D3Q19").shape == (19, 4)
    finally:
        capi._lib = default
    assert bfa.load_library() is default


def test_abi_version_and_counts(bfa):
    lib = bfa.load_library()
    assert lib.bf_abi_version() == 1
    assert [lib.bf_velocity_count(v) for v in (2, 3, 4, 5)] == [9, 15, 19, 27]
    assert lib.bf_velocity_count(7) == 0


@pytest.mark.parametrize("name", ["D2Q9", "D3Q15", "D3Q19", "D3Q27"])
def test_product_tables_equal_oracle(bfa, orc, name):
    a = bfa.velocity_table(name)
    b = orc.velocity_table(name)
    np.testing.assert_array_equal(a[:, :3], b[:, :3])
    np.testing.assert_array_equal(a[:, 3], b[:, 3].astype(np.float32))


@pytest.mark.parametrize("name", ["D2Q5", "D3Q7"])
def test_product_thermal_tables_equal_oracle(bfa, orc, name):
    a = bfa.velocity_table(name, thermal=True)
    b = orc.velocity_table(name, thermal=True)
    np.testing.assert_array_equal(a[:, :3], b[:, :3])
    np.testing.assert_array_equal(a[:, 3], b[:, 3].astype(np.float32))


@pytest.mark.parametrize("name,thermal", [("D2Q9", False), ("D3Q15", False), ("D3Q19", False), ("D3Q27", False),
                                          ("D2Q5", True), ("D3Q7", True)])
def test_rest_population_mass_fix(orc, name, thermal):
    """The FP32 weights plus the rest-population correction sum to exactly 1;
    the correction is a power of two (zero for D3Q7), and without it the FP32
    weights do not sum to 1 (the case the correction exists for)."""
    w = orc.velocity_table(name, thermal=thermal)[:, 3].astype(np.float32)
    vid = {"D2Q5": 0, "D3Q7": 1}[name] if thermal else orc.VELOCITY_SETS[name]
    fix = np.float32(orc.load("f32").orc_rest_fix_f32(vid, 1 if thermal else 0))
    total = sum(Fraction(float(v)) for v in w)
    assert total + Fraction(float(fix)) == 1
    if name == "D3Q7":
        assert fix == 0
    else:
        assert total != 1 and fix < 0 and np.log2(-float(fix)) == int(np.log2(-float(fix)))


def test_bits_per_boundary_known_answers(bfa):
    """The library's rule against the known answers of LBM.cpp:569 (worked by
    hand per range of counts in test_oracle.expected_bits), not against the
    oracle's restatement of the same formula."""
    from test_oracle import expected_bits
    for n in range(0, 300):
        assert bfa.bits_per_boundary(n) == expected_bits(n), n


@pytest.mark.parametrize("bits", [1, 2, 4, 8])
def test_pack_matches_oracle(bfa, orc, bits):
    rng = np.random.default_rng(7)
    n = 777
    ids = rng.integers(0, 1 << bits, n).astype(np.uint8)
    words = bfa.pack_boundaries(ids, bits)
    ref = np.zeros_like(words)
    orc.load().orc_pack(ids.ctypes.data_as(C.POINTER(C.c_uint8)), n, bits, ref.ctypes.data_as(C.POINTER(C.c_uint32)))
    np.testing.assert_array_equal(words, ref)
    np.testing.assert_array_equal(bfa.unpack_boundaries(words, n, bits), ids)


def test_pack_rejects_oversized_ids(bfa):
    with pytest.raises(bfa.BFError):
        bfa.pack_boundaries(np.array([0, 2], np.uint8), 1)


@pytest.mark.parametrize("kw,msg", [
    (dict(resolution=(0, 4, 4)), "resolution"),
    (dict(resolution=(8, 8, 2), velocity_set="D2Q9"), "dimension"),
    (dict(relaxation_time=0.5), "relaxation"),
    (dict(storage=7), "storage"),
    (dict(velocity_set=9), "velocity set"),
    (dict(boundary_mode=5), "boundary mode"),
])
def test_create_rejects_bad_config_before_device(bfa, kw, msg):
    args = dict(resolution=(8, 8, 8), relaxation_time=0.6, velocity_set="D3Q19")
    args.update(kw)
    with pytest.raises(bfa.BFError) as e:
        bfa.Solver(**args)
    assert e.value.status == 1
    assert msg in str(e.value)


def test_null_handle_is_rejected(bfa):
    lib = bfa.load_library()
    assert lib.bf_step(None, 1, 0) == 1
    assert lib.bf_ticks(None) == -1
    lib.bf_destroy(None)



def test_field_copy_rejects_null_handles(bfa):
    """bf_copy_field_to_device checks its arguments before any device call (no
    GPU needed): a null handle or destination is BF_ERR_INVALID (1)."""
    lib = bfa.load_library()
    assert lib.bf_copy_field_to_device(None, 0, None, 16) == 1


def test_probes_live_outside_the_solver_library(bfa):
    """The bandwidth probes are the benchmark's (libbfprobe.so); the solver
    library exports none of them."""
    lib = bfa.load_library()
    for name in ("bf_probe_copy_bandwidth", "bf_probe_aa_bandwidth", "bfp_copy_bandwidth", "bfp_aa_bandwidth"):
        assert not hasattr(lib, name), name


def test_aa_probe_rejects_bad_arguments():
    """bfp_aa_bandwidth checks its arguments before any device call: rows
    longer than one block (X > 1024), empty grids and a null result are
    BF_ERR_INVALID (1)."""
    from boltzmannfluids_amd import probe
    lib = probe.load_library()
    out = C.c_double()
    assert lib.bfp_aa_bandwidth(0, 2048, 8, 8, 1, 1, C.byref(out)) == 1
    assert lib.bfp_aa_bandwidth(0, 0, 8, 8, 0, 1, C.byref(out)) == 1
    assert lib.bfp_aa_bandwidth(0, 64, 8, 8, 0, 1, None) == 1


def test_pull_probe_rejects_bad_arguments():
    """bfp_pull_bandwidth checks its arguments before any device call, as the
    AA probe does."""
    from boltzmannfluids_amd import probe
    lib = probe.load_library()
    out = C.c_double()
    assert lib.bfp_pull_bandwidth(0, 2048, 8, 8, 1, C.byref(out)) == 1
    assert lib.bfp_pull_bandwidth(0, 8, 0, 8, 1, C.byref(out)) == 1
    assert lib.bfp_pull_bandwidth(0, 64, 8, 8, 0, C.byref(out)) == 1
    assert lib.bfp_pull_bandwidth(0, 64, 8, 8, 1, None) == 1
