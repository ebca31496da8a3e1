"""Host-side checks of the product library that need no GPU: the C ABI loads and exports
every declared symbol, argument errors are reported before any device work, and the host
parts (training, statistics, the chi-square LUT, statistics files) match the oracle."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def D():
    import despeckcl_amd
    despeckcl_amd.lib()
    return despeckcl_amd


def _declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if not f.endswith(".h"):
            continue
        txt = open(os.path.join(inc, f)).read()
        if 'extern "C"' not in txt:
            continue
        names |= set(re.findall(r"\b(dkx_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol(D):
    lib = D.lib()
    names = _declared_symbols()
    assert len(names) >= 13
    for n in names:
        assert hasattr(lib, n), n
    from despeckcl_amd._lib import SIGNATURES
    assert set(SIGNATURES) == set(names)


def test_cxx_api_symbols_exported(D):
    syms = os.popen(f"nm -D --defined-only {D._lib.LIB_PATH}").read()
    for mangled_part in ("despeckcl5nlsar", "despeckcl7nlinsar", "despeckcl14nlsar_training",
                         "store_nlsar_stats_collection", "load_nlsar_stats_collection"):
        assert mangled_part in syms, mangled_part


def _stats_for(D, keys):
    coll = D.nlsar_stats_collection()
    for p, s in keys:
        st = D.stats(lut_size=256)
        st.dissims_min, st.dissims_max = 1.0, 10.0
        st.quantilles = np.linspace(0, 0.99, 256).astype(np.float32)
        st.chi2cdf_inv = np.linspace(0, 100, 256).astype(np.float32)
        coll[D.params(p, s)] = st
    return coll


def test_argument_errors_without_device(D):
    a = np.ones((16, 16), np.float32)
    st = _stats_for(D, [(3, 1), (5, 1)])
    with pytest.raises(ValueError, match="odd"):
        D.nlsar_insar(a, a, a, 20, [3, 5], [1], st)
    with pytest.raises(ValueError, match="every .patch size, scale size. pair"):
        D.nlsar_insar(a, a, a, 21, [3, 5, 7], [1], st)
    with pytest.raises(ValueError, match=r"no statistics for \(patch 7, scale 1\)"):
        D.nlsar_insar(a, a, a, 21, [3, 7], [1], st)
    with pytest.raises(ValueError, match="odd"):
        D.nlsar_insar(a, a, a, 21, [3, 5], [2], _stats_for(D, [(3, 2), (5, 2)]))
    with pytest.raises(ValueError, match="shape"):
        D.nlsar_insar(a, a[:8], a, 21, [3, 5], [1], st)
    with pytest.raises(ValueError, match="empty"):
        z = np.zeros((0, 5), np.float32)
        D.nlsar_insar(z, z, z, 21, [3, 5], [1], st)
    bad = _stats_for(D, [(3, 1), (5, 1)])
    bad[D.params(3, 1)].quantilles[4] = 1.5
    with pytest.raises(ValueError, match="quantile"):
        D.nlsar_insar(a, a, a, 21, [3, 5], [1], bad)
    for d in (1, 4):
        c = np.ones((d, d, 16, 16), np.complex64)
        with pytest.raises(ValueError, match="dimension 2 or 3"):
            D.nlsar(c, 21, [3, 5], [1], st)
    with pytest.raises(ValueError, match="square"):
        D.nlsar(np.ones((2, 3, 16, 16), np.complex64), 21, [3, 5], [1], st)
    with pytest.raises(ValueError, match="odd"):
        D.nlsar(np.ones((3, 3, 16, 16), np.complex64), 20, [3, 5], [1], st)
    with pytest.raises(ValueError, match="odd"):
        D.boxcar(a, a, a, 4)
    with pytest.raises(ValueError, match="above 15"):
        D.boxcar(a, a, a, 17)
    with pytest.raises(ValueError, match="patch sizes 16, 32 and 64"):
        D.goldstein(a, a, a, 48, 4, 0.5)
    with pytest.raises(ValueError, match="overlap"):
        D.goldstein(a, a, a, 32, 16, 0.5)
    with pytest.raises(ValueError, match="quarter"):
        D.goldstein(a, a, a, 16, 5, 0.5)
    with pytest.raises(ValueError, match="odd, up to 21"):
        D.nlinsar(a, a, a, 21, 23, 2, 10)
    with pytest.raises(ValueError, match="odd"):
        D.nlinsar(a, a, a, 21, 4, 2, 10)
    with pytest.raises(ValueError, match="odd"):
        D.nlinsar(a, a, a, 22, 7, 2, 10)
    with pytest.raises(ValueError, match="lmin"):
        D.nlinsar(a, a, a, 21, 7, 2, 0)


def test_images_beyond_32_bit_grid_offsets_are_rejected(D):
    """The kernels address their padded grids with 32-bit offsets: images beyond that are
    rejected with a message to split them into bands (the reference tiles any size on the
    host), before any device call.  The pointers are never dereferenced."""
    from despeckcl_amd import device as DV
    st = _stats_for(D, [(3, 1), (5, 1), (7, 1), (3, 3), (5, 3), (7, 3)])
    for n in (20000, 40000):
        with pytest.raises(ValueError, match="too large for one call"):
            DV.nlsar_insar(8, 8, 8, 8, 8, 8, n, n, 21, [3, 5, 7], [1, 3], st)
    with pytest.raises(ValueError, match="too large for one call"):
        DV.nlinsar(8, 8, 8, 8, 8, 8, 40000, 40000, 21, 7, 2, 10)


def test_statistics_collections_with_extra_keys(D):
    """A collection trained on more (patch, scale) keys than the call filters with passes the
    host checks, as in the reference (only the device step fails here, without a GPU); a
    missing key or a key given twice is rejected before any device call."""
    from despeckcl_amd import _lib as L
    a = np.ones((16, 16), np.float32)
    sup = _stats_for(D, [(3, 1), (5, 1), (7, 1), (3, 3)])
    with pytest.raises(Exception) as e:
        D.nlsar_insar(a, a, a, 21, [3, 5], [1], sup)
    assert not isinstance(e.value, ValueError), e.value      # not a statistics error
    arr, n, keep = D._stats_structs(_stats_for(D, [(3, 1), (5, 1)]))
    dup = (L.DkxStats * 3)(arr[0], arr[1], arr[1])
    P, nP = L.int_array([3, 5])
    S_, nS = L.int_array([1])
    out = [np.zeros((16, 16), np.float32) for _ in range(3)]
    rc = L.lib().dkx_nlsar_insar(a.ctypes.data, a.ctypes.data, a.ctypes.data, out[0].ctypes.data,
                                 out[1].ctypes.data, out[2].ctypes.data, 16, 16, 21, P, nP, S_, nS,
                                 dup, 3, 15.0, 49.0)
    assert rc == L.DKX_ESTATS and b"twice" in L.lib().dkx_last_error()


def test_stats_file_oversized_lut_is_rejected(D, tmp_path):
    """A malformed statistics file whose LUT size token would overflow the length check is
    an ordinary error, not an abort (no C++ exception crosses the C ABI)."""
    f = tmp_path / "huge.txt"
    f.write_text("despeckcl_amd-nlsar-stats 1 1 3 1 18446744073709551615 1.0 2.0 0.5\n")
    with pytest.raises(ValueError, match="not a readable"):
        D.load_nlsar_stats_collection(str(f))
    f.write_text("despeckcl_amd-nlsar-stats 1 1 3 1 4000000000 1.0 2.0 0.5\n")
    with pytest.raises(ValueError, match="not a readable"):
        D.load_nlsar_stats_collection(str(f))


def test_tile_settings(D):
    """Tile shapes: both filters default to the automatic choice (-1) and accept -1..2; the
    setting round-trips and invalid values are rejected."""
    from despeckcl_amd import device as DV
    assert DV.get_tile(1) == -1
    assert DV.get_tile(0) == -1
    for f in (0, 1):
        for t in (2, 0, 1, -1):
            DV.set_tile(f, t)
            assert DV.get_tile(f) == t
    for f, t in ((0, -2), (0, 3), (1, -2), (1, 3), (2, 0)):
        with pytest.raises(ValueError, match="tile"):
            DV.set_tile(f, t)
    assert DV.get_tile(0) == -1 and DV.get_tile(1) == -1


def test_tuned_tile_records_without_device(D):
    """The autotuner's record lookup needs no device: nothing is recorded in a fresh process
    (or after clearing), and invalid filters or sizes read as no record."""
    from despeckcl_amd import device as DV
    DV.clear_tuned_tiles()
    assert DV.tuned_tile(0, 2048, 2048, 21, [3, 5, 7], [1, 3]) == -1
    assert DV.tuned_tile(0, 2048, 2048, 21, [3, 5, 7], [1, 3], every_key=True) == -1
    assert DV.tuned_tile(1, 4096, 4096, 21, [7]) == -1
    assert DV.tuned_tile(2, 64, 64, 21, [7]) == -1
    assert DV.tuned_tile(1, 0, 64, 21, [7]) == -1


def test_nlinsar_simulated_h_reproduces_the_tabulated_values(D):
    """quantile_insar (insarsim_simu.cpp:62-82) restated with a seeded generator.  The reference
    tabulates its result for patch 3 (0.58) and 7 (0.244) at alpha 0.92 (nlinsar.cpp:65-76):
    the simulation must give them within its Monte-Carlo spread, per seed (6 %) and on the mean
    of four seeds (2 %).  A wrong quantile, a missing negation, the patch side instead of its
    area, or correlated pixels fail these bounds.  The value decreases with the patch size and
    depends on the seed only."""
    from despeckcl_amd import device as DV
    for p, want in ((3, 0.58), (7, 0.244)):
        vals = [DV.nlinsar_simulate_h(p, seed) for seed in (5489, 1, 2, 3)]
        print(p, vals)
        assert all(abs(v - want) <= 0.06 * want for v in vals), (p, vals)
        assert abs(np.mean(vals) - want) <= 0.02 * want, (p, vals)
    hs = [DV.nlinsar_simulate_h(p) for p in (1, 3, 5, 7, 9, 11)]
    assert all(a > b for a, b in zip(hs, hs[1:])), hs
    assert DV.nlinsar_simulate_h(5, 7) == DV.nlinsar_simulate_h(5, 7)
    assert DV.nlinsar_simulate_h(5, 7) != DV.nlinsar_simulate_h(5, 8)
    for bad in (0, 2, 23):
        with pytest.raises(ValueError, match="odd"):
            DV.nlinsar_simulate_h(bad)


def test_training_needs_a_device(D):
    """Training runs on the GPU only: without a device it fails loudly (arguments are still
    checked first)."""
    from despeckcl_amd import hip
    from despeckcl_amd import synthetic as S
    a1, a2, ph = S.homogeneous_area(16, 1)
    with pytest.raises(ValueError, match="odd"):
        D.nlsar_train_insar(a1, a2, ph, [4], [1])
    if hip.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(D.DespeckleError, match="no HIP device"):
        D.nlsar_train_insar(a1, a2, ph, [3], [1])


def test_stats_constructor_matches_oracle(D, oracle):
    rng = np.random.default_rng(4)
    d = rng.gamma(2.0, 2.0, 5000).astype(np.float32)
    d[::97] = np.nan
    s = D.stats(d, 256)
    o, _ = oracle.stats_build(d, 256)
    assert s.dissims_min == o.dissims_min and s.dissims_max == o.dissims_max
    assert np.array_equal(s.quantilles, o.quantilles)
    assert np.array_equal(s.chi2cdf_inv, o.chi2cdf_inv)


def test_chi2_matches_scipy_fixture(D):
    ref = np.load(os.path.join(GOLDEN, "chi2_49_lut256.npy"))
    got = np.array([np.float32(D.lib().dkx_chi2_inv(i / 256.0, 49.0)) for i in range(255)])
    assert np.array_equal(got, ref)


def test_stats_file_roundtrip(D, tmp_path):
    fx = D.load_nlsar_stats_collection(os.path.join(GOLDEN, "train32_stats.txt"))
    assert sorted(fx.keys()) == [(3, 1), (3, 3), (5, 1), (5, 3), (7, 1), (7, 3)]
    f = tmp_path / "s.txt"
    D.store_nlsar_stats_collection(fx, str(f))
    back = D.load_nlsar_stats_collection(str(f))
    for k in fx:
        assert back[k].dissims_min == fx[k].dissims_min
        assert np.array_equal(back[k].quantilles, fx[k].quantilles)
        assert np.array_equal(back[k].chi2cdf_inv, fx[k].chi2cdf_inv)


def test_stats_file_is_a_boost_text_archive(D, tmp_path):
    """store_nlsar_stats_collection writes the reference's format: Boost's text archive of
    std::map<nlsar::params, nlsar::stats> (nlsar_training.cpp:99-104; members as serialized in
    parameters.h:48-55 and stats.h:40-50)."""
    fx = D.load_nlsar_stats_collection(os.path.join(GOLDEN, "train32_stats.txt"))
    f = tmp_path / "nlsar_stats.txt"
    D.store_nlsar_stats_collection(fx, str(f))
    tok = f.read_text().split()
    # signature, library version, map class info, count, item version, class info of the
    # pair and of params, the first key, class info of stats, lut size
    assert tok[:16] == ["22", "serialization::archive", "17", "0", "0", "6", "0", "0", "0", "0", "0",
                        "3", "1", "0", "0", "256"]
    assert float(tok[16]) == np.float32(fx[(3, 1)].dissims_min)
    assert tok[18:20] == ["256", "0"]          # quantilles: size, item version
    assert np.array_equal(np.array(tok[20:276], np.float32), fx[(3, 1)].quantilles)
    assert tok[276:278] == ["256", "0"]
    assert tok[534:536] == ["3", "3"]          # second key, no class info any more
    assert len(tok) == 3 + 2 + 2 + 6 + 6 * (2 + 3 + 2 * (2 + 256))


def test_load_boost_archive_layouts(D, tmp_path):
    """Archives laid out as Boost writes them load: the current layout, 8-digit fixed-point
    floats of older Boost versions, and library version 3 (no item versions).  No Boost is
    available here to produce a real file, so the byte-exact output of a given Boost
    version is parity unpinned; the layout follows the serialize() members above."""
    e1 = "3 1 {si}3 1.5 2.25 3 {iv}0 0.5 0.75 3 {iv}1 2 100000"
    e2 = "5 1 3 1.25 3.5 3 {iv}0 0.25 0.5 3 {iv}4 5 6"
    cases = {
        "v17": "22 serialization::archive 17 0 0 2 0 0 0 0 0 " + e1.format(si="0 0 ", iv="0 ") + " " + e2.format(iv="0 "),
        "v12 fixed": "22 serialization::archive 12 0 0 2 0 0 0 0 0 " + e1.format(si="0 0 ", iv="0 ") + " "
                     + e2.format(iv="0 ").replace("1.25", "1.2500000"),
        "v3": "22 serialization::archive 3 0 0 2 0 0 0 0 " + e1.format(si="0 0 ", iv="") + " " + e2.format(iv=""),
    }
    for label, text in cases.items():
        f = tmp_path / "a.txt"
        f.write_text(text + "\n")
        got = D.load_nlsar_stats_collection(str(f))
        assert sorted(got.keys()) == [(3, 1), (5, 1)], label
        a, b = got[(3, 1)], got[(5, 1)]
        assert a.lut_size == 3 and a.dissims_min == 1.5 and a.dissims_max == 2.25, label
        assert list(a.quantilles) == [0, 0.5, 0.75] and list(a.chi2cdf_inv) == [1, 2, 100000], label
        assert b.dissims_min == 1.25 and list(b.chi2cdf_inv) == [4, 5, 6], label
    bad = tmp_path / "bad.txt"
    bad.write_text(cases["v17"].rsplit(" ", 1)[0] + "\n")     # one value short
    with pytest.raises(ValueError, match="statistics file"):
        D.load_nlsar_stats_collection(str(bad))
    with pytest.raises(ValueError, match="cannot open"):
        D.load_nlsar_stats_collection(str(tmp_path / "missing.txt"))


CXX_ROUNDTRIP = r"""
#include "despeckcl.h"
#include <cstdio>
#include <stdexcept>
int main(int argc, char** argv)
{
    nlsar_stats_collection a = despeckcl::load_nlsar_stats_collection(argv[1]);
    despeckcl::store_nlsar_stats_collection(a, argv[2]);
    nlsar_stats_collection b = despeckcl::load_nlsar_stats_collection(argv[2]);
    int same = a.size() == b.size();
    for (const auto& kv : a) {
        const nlsar::stats& x = kv.second;
        const nlsar::stats& y = b.at(kv.first);
        same &= x.lut_size == y.lut_size && x.dissims_min == y.dissims_min &&
                x.dissims_max == y.dissims_max && x.quantilles == y.quantilles &&
                x.chi2cdf_inv == y.chi2cdf_inv;
    }
    try {
        despeckcl::load_nlsar_stats_collection(argv[3]);
        same = 0;
    } catch (const std::exception&) {
    }
    std::printf("%zu %d\n", a.size(), same);
    return 0;
}
"""


def test_cxx_api_stats_files_roundtrip(D, tmp_path):
    """A program written against despeckcl.h stores and loads statistics through the C++ API
    (nlsar_training.cpp:99-113), and a missing file throws."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "rt.cpp"
    src.write_text(CXX_ROUNDTRIP)
    exe = tmp_path / "rt"
    libdir = os.path.dirname(D._lib.LIB_PATH)
    subprocess.run([gxx, "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src),
                    "-L", libdir, "-ldespeckcl_amd", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([str(exe), os.path.join(GOLDEN, "train32_stats.txt"), str(tmp_path / "s.txt"),
                          str(tmp_path / "missing.txt")], capture_output=True, text=True, check=True).stdout
    assert out.split() == ["6", "1"]
    assert (tmp_path / "s.txt").read_text().startswith("22 serialization::archive 17 ")


def test_synthetic_is_seeded():
    from despeckcl_amd import synthetic as S
    a = S.interferogram(64, 48, 5)
    b = S.interferogram(64, 48, 5)
    c = S.interferogram(64, 48, 6)
    for x, y in zip(a, b):
        assert x.dtype == np.float32 and np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    assert np.all(a[0] >= 0) and np.all(np.abs(a[2]) <= np.pi)


def test_no_oracle_in_product():
    """The product package never imports or loads the oracle."""
    pkg = os.path.join(ROOT, "despeckcl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.replace("# oracle", ""), f


LUT_INDEX_C = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static float bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t ubits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
/* the scale pass's index: q0 = a * (1/rng), e = fma(-q0, rng, a), q = fma(e, 1/rng, q0) */
static unsigned idx_fma(float a, float rng, float r, float scl) {
    float q0 = a * r, e = fmaf(-q0, rng, a), q = fmaf(e, r, q0);
    return (unsigned)(q * scl);
}
/* the reference's: (unsigned int)((d - min) / (max - min) * (lut_size - 1)) */
static unsigned idx_div(float a, float rng, float scl) { return (unsigned)(a / rng * scl); }
int main(void) {
    const float dmins[] = {0.0f, -3.25f, 1.5f, 12.0f, -0.001f, 100.0f};
    const float dmaxs[] = {1.0f, 7.3421f, 255.0f, 164.77f, 0.099f, 98865.4f};
    long long n = 0, bad = 0;
    uint64_t s = 88172645463325252ull;
    for (int k = 0; k < 6; k++) {
        const float dmin = dmins[k], dmax = dmaxs[k], rng = dmax - dmin, r = 1.0f / rng;
        for (int lut = 2; lut <= 1024; lut *= 2) {
            const float scl = (float)(lut - 1);
            /* +-64 ulps around every index boundary i * rng / scl */
            for (int i = 0; i <= lut; i++) {
                const float b = (float)i * rng / scl;
                const uint32_t c = ubits(b);
                for (int dd = -64; dd <= 64; dd++) {
                    if (c < 64 && dd < 0 && (uint32_t)(-dd) > c) continue;
                    const float a = bits(c + dd);
                    if (!(a >= 0.0f && a <= rng)) continue;
                    n++;
                    if (idx_fma(a, rng, r, scl) != idx_div(a, rng, scl)) bad++;
                }
            }
            /* random differences in [0, rng] */
            for (int t = 0; t < 200000; t++) {
                s ^= s << 13; s ^= s >> 7; s ^= s << 17;
                const float d = dmin + (float)((s >> 11) * (1.0 / 9007199254740992.0)) * rng;
                const float dc = fminf(fmaxf(d, dmin), dmax), a = dc - dmin;
                n++;
                if (idx_fma(a, rng, r, scl) != idx_div(a, rng, scl)) bad++;
            }
        }
    }
    printf("%lld %lld\n", n, bad);
    return 0;
}
"""


def test_lut_index_fma_quotient_equals_division(tmp_path):
    """The scale pass forms the LUT index from the FMA-corrected quotient (Markstein) instead
    of a division; it must give the reference expression's index for every difference: at
    +-64 ulps around every index boundary and on random differences, for several trained
    ranges and LUT sizes.  (An exhaustive run over all floats of ten ranges agrees too.)"""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "lutidx.c"
    src.write_text(LUT_INDEX_C)
    exe = tmp_path / "lutidx"
    subprocess.run([gcc, "-O2", "-march=x86-64-v3", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"],
                   check=True)
    n, bad = map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split())
    assert n > 1_000_000
    assert bad == 0
