"""The tuning table that replaces the reference's compile-time tuning headers
(wjc404/GEMM_AVX2 dgemm_tune.h:6-10, sgemm_tune.h:7-11, clamped in
src/gemm_set_parameters.h:10-61): on the CPU, the libraries' reader and the
autotuner's writer agree, the same table always gives the same kernel choice
(determinism), forced configurations and format checks behave as documented.
No device is needed: *_x_tune_lookup only consults the table."""
import os

import pytest

import gemm_avx2_amd as gx
from gemm_avx2_amd import autotune as at

DEFAULT = os.path.join(gx.LIB_DIR, "gemm_mi355x_tune.txt")


@pytest.fixture
def restore_tables():
    yield
    for d in ("d", "s"):
        gx.set_tile_config(d, -1)
        gx.load_tuning(DEFAULT if os.path.exists(DEFAULT) else None, d)


def _bucket_edge(b, scale):
    """Lowest (scale 0), central (1) and highest (2) dimension of bucket b:
    bucket b holds [2^b/sqrt(2), 2^b*sqrt(2)), so its lowest member is
    ceil(2^(b-1/2)) = isqrt(2^(2b-1) - 1) + 1."""
    import math
    lowest = lambda c: math.isqrt((1 << (2 * c - 1)) - 1) + 1
    return (lowest(b), 1 << b, lowest(b + 1) - 1)[scale]


def _entries():
    return {("f64", "NN", 13, 13, 13): 6, ("f64", "TN", 8, 8, 8): 1, ("f64", "NT", 10, 12, 9): 3,
            ("f32", "NN", 13, 13, 13): 2, ("f32", "TT", 4, 15, 15): 5}


def test_write_read_roundtrip_is_deterministic(tmp_path):
    p1, p2 = str(tmp_path / "a.txt"), str(tmp_path / "b.txt")
    e = _entries()
    notes = {k: f"note {i}" for i, k in enumerate(e)}
    at.write_table(p1, e, notes, "gfx950", ["header"])
    # insertion order must not matter
    at.write_table(p2, dict(reversed(list(e.items()))), notes, "gfx950", ["header"])
    assert open(p1).read() == open(p2).read()
    got, got_notes, arch, rocm = at.read_table(p1)
    assert got == e and arch == "gfx950" and got_notes == notes and rocm == gx.build_rocm()


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_library_lookup_follows_the_table(tmp_path, dtype, restore_tables):
    path = str(tmp_path / "t.txt")
    e = _entries()
    at.write_table(path, e, {}, None)
    assert gx.load_tuning(path, dtype) == len(e)
    want = "f64" if dtype == "d" else "f32"
    for (dt, tr, mb, nb, kb), cfg in e.items():
        if dt != want:
            continue
        for scale in (0, 1, 2):  # lowest, middle and highest size in each bucket
            m, n, k = (_bucket_edge(b, scale) for b in (mb, nb, kb))
            if mb == 4:
                m = 3  # buckets clamp: anything below 23 is bucket 4
            assert gx.tune_lookup(dtype, tr[0], tr[1], m, n, k) == cfg
            # only the first character counts, and lower case is the same
            assert gx.tune_lookup(dtype, tr[0].lower(), "C" if tr[1] == "T" else "n",
                                  m, n, k) == cfg
            # the same table always gives the same choice
            assert gx.tune_lookup(dtype, tr[0], tr[1], m, n, k) == cfg
    # outside the table: the size heuristic (128x128 tiles when there are
    # enough of them for every CU, else 64x64)
    assert gx.tune_lookup(dtype, "N", "N", 8192, 8192, 64) == 0
    assert gx.tune_lookup(dtype, "N", "N", 100, 100, 100) == 1
    assert gx.tune_lookup(dtype, "N", "N", -1, 1, 1) == -1


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_forced_configuration_overrides_the_table(tmp_path, dtype, restore_tables):
    path = str(tmp_path / "t.txt")
    at.write_table(path, _entries(), {}, None)
    gx.load_tuning(path, dtype)
    gx.set_tile_config(dtype, 4)
    assert gx.tune_lookup(dtype, "N", "N", 8192, 8192, 8192) == 4
    assert gx.tune_lookup(dtype, "T", "N", 300, 300, 300) == 4
    gx.set_tile_config(dtype, -1)
    assert gx.tune_lookup(dtype, "N", "N", 8192, 8192, 8192) == (6 if dtype == "d" else 2)
    with pytest.raises(gx.GemmError):
        gx.set_tile_config(dtype, gx.TILE_CONFIGS)


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_split_ranges_follow_the_kernels_slice_depth(dtype):
    """The K ranges and partial-tile bytes the host reserves for a forced
    configuration and split are those the kernels run: ranges of
    ceil(ceil(k / split) / BK) BK-deep slices with the BK of that element
    type's kernels (configuration 7 takes 16-deep FP64 and 32-deep FP32
    slices; sizing the FP64 workspace with 32 gave 3 ranges where the kernel
    runs 4 at k = 160, split 4).  Needs no device."""
    es = 8 if dtype == "d" else 4
    try:
        for cfg in at.SPLIT_CONFIGS:
            gx.set_tile_config(dtype, cfg)
            bk = at.TILE_BK[dtype][cfg]
            bm, bn = at.TILE_DIMS[cfg]
            for split in (2, 3, 4, 8, 16):
                gx.set_split_config(dtype, split)
                for k in (16, 40, 100, 144, 160, 176, 192, 600, 640, 1040, 4111):
                    kc = -(-(-(-k // split)) // bk) * bk
                    ns = -(-k // kc)
                    m, n = 200, 130
                    tiles = -(-m // bm) * -(-n // bn)
                    got = gx.tune_lookup_split(dtype, "N", "N", m, n, k)
                    want = (ns, ns * tiles * bm * bn * es if ns > 1 else 0)
                    assert got == want, (cfg, split, k, got, want)
    finally:
        gx.set_tile_config(dtype, -1)
        gx.set_split_config(dtype, -1)


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_reader_rejects_other_formats_and_keeps_the_previous_table(tmp_path, dtype,
                                                                    restore_tables):
    good = str(tmp_path / "good.txt")
    at.write_table(good, _entries(), {}, None)
    assert gx.load_tuning(good, dtype) == 5
    bad_format = tmp_path / "bad1.txt"
    bad_format.write_text("format 99\nf64 NN 13 13 13 1\nf32 NN 13 13 13 1\n")
    no_format = tmp_path / "bad2.txt"
    no_format.write_text("f64 NN 13 13 13 1\nf32 NN 13 13 13 1\n")
    for bad in (bad_format, no_format, tmp_path / "missing.txt"):
        assert gx.load_tuning(str(bad), dtype) == -1
        assert gx.tune_lookup(dtype, "N", "N", 8192, 8192, 8192) == (6 if dtype == "d" else 2)
    with pytest.raises(ValueError):
        at.read_table(str(bad_format))
    # invalid lines are skipped, a later entry replaces an earlier one
    mixed = tmp_path / "mixed.txt"
    mixed.write_text(f"format 2\nrocm {gx.build_rocm()}\n# comment\n"
                     "f64 NN 13 13 13 2\nf64 NN 13 13 13 5  # later\n"
                     "f32 NN 13 13 13 2\nf32 NN 13 13 13 5\n"
                     "f64 NN 3 13 13 1\nf64 NN 13 13 13 99\nf16 NN 13 13 13 1\n"
                     "f32 NN 3 13 13 1\nf32 NN 13 13 13 99\n")
    assert gx.load_tuning(str(mixed), dtype) == 4
    assert gx.tune_lookup(dtype, "N", "N", 8192, 8192, 8192) == 5
    py, _, _, _ = at.read_table(str(mixed))
    assert py[("f64", "NN", 13, 13, 13)] == 5
    assert gx.load_tuning(None, dtype) == 0
    assert gx.tune_lookup(dtype, "N", "N", 8192, 8192, 8192) == 0


def test_resume_skips_measured_keys():
    wanted = [("d", "NN", 8192, 8192, 8192), ("d", "NN", 9000, 9000, 9000),
              ("d", "TN", 256, 256, 256), ("s", "NN", 8192, 8192, 8192)]
    have = {at.key_of("d", "NN", 8192, 8192, 8192): 0}
    # one shape per key (9000 falls in the 8192 bucket: 2^13 nearest on a log scale)
    assert at.pending(wanted, {}, False) == [wanted[0], wanted[2], wanted[3]]
    assert at.pending(wanted, have, True) == [wanted[2], wanted[3]]
    assert at.bucket(1) == 4 and at.bucket(16) == 4 and at.bucket(22) == 4
    assert at.bucket(23) == 5 and at.bucket(5792) == 12 and at.bucket(5793) == 13
    assert at.bucket(8191) == 13 and at.bucket(11585) == 13 and at.bucket(11586) == 14
    assert at.bucket(1 << 20) == 15 and at.bucket(0) == 4


def test_configuration_counts_agree():
    import re
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "include", "gemm_mi355x.h")).read()
    n = int(re.search(r"#define GEMMX_TILE_CONFIGS (\d+)", hdr).group(1))
    nb = int(re.search(r"#define GEMMX_BATCH_CONFIGS (\d+)", hdr).group(1))
    assert gx.TILE_CONFIGS == at.TILE_CONFIGS == n
    assert gx.BATCH_CONFIGS == nb
    # the autotuner's copy of each configuration's tile, slice depths and
    # split-K ability follows the kernels' own table (kNativeCfg: bm, bn,
    # bk64, bk32, split_ok)
    impl = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                             "gemm_avx2_amd", "csrc", "gemm_native_impl.h")).read()
    body = re.search(r"kNativeCfg\[GEMMX_NCFG_COUNT\] = \{(.*?)\};", impl, re.S).group(1)
    cfgs = [tuple(int(v) for v in t)
            for t in re.findall(r"\{(\d+), (\d+), (\d+), (\d+), (\d)\}", body)]
    assert len(cfgs) == n
    assert list(at.TILE_DIMS) == [(bm, bn) for bm, bn, _, _, _ in cfgs]
    assert list(at.TILE_BK["d"]) == [t[2] for t in cfgs]
    assert list(at.TILE_BK["s"]) == [t[3] for t in cfgs]
    assert tuple(at.SPLIT_CONFIGS) == tuple(c for c, t in enumerate(cfgs) if t[4])


def test_emulated_entries_pick_the_dgemm_arithmetic(tmp_path, restore_tables):
    """`emu` lines (written by the autotuner) choose dgemm_'s arithmetic in
    the automatic mode per shape bucket; the same table always gives the same
    choice, invalid counts are skipped, and the Python reader agrees."""
    path = str(tmp_path / "t.txt")
    e = dict(_entries())
    e[("emu", "NN", 13, 13, 13)] = ("ozaki2", 18)
    e[("emu", "TN", 12, 12, 12)] = ("native", 0)
    e[("emu", "NT", 14, 14, 10)] = ("ozaki", 6)
    at.write_table(path, e, {}, None)
    got, _, _, _ = at.read_table(path)
    assert got == e
    assert gx.load_tuning(path, "d") == len(e)
    assert gx.tune_lookup_emulated("N", "N", 8192, 9000, 11000) == (gx.MODE_OZAKI2, 18)
    assert gx.tune_lookup_emulated("n", "n", 8192, 9000, 11000) == (gx.MODE_OZAKI2, 18)
    assert gx.tune_lookup_emulated("T", "N", 4096, 4096, 4096) == (gx.MODE_NATIVE, 0)
    assert gx.tune_lookup_emulated("N", "T", 16384, 16384, 1024) == (gx.MODE_OZAKI, 6)
    assert gx.tune_lookup_emulated("T", "T", 8192, 8192, 8192) == (gx.MODE_NATIVE, 0)
    # the native tile entries are unaffected
    assert gx.tune_lookup("d", "N", "N", 8192, 8192, 8192) == 6
    bad = tmp_path / "bad.txt"
    bad.write_text(f"format 2\nrocm {gx.build_rocm()}\nemu NN 13 13 13 ozaki2 40\n"
                   "emu NN 13 13 13 ozaki 2\n"
                   "emu NN 13 13 13 fp16 4\nemu NN 13 13 13 ozaki2 20\n")
    assert gx.load_tuning(str(bad), "d") == 1
    assert gx.tune_lookup_emulated("N", "N", 8192, 8192, 8192) == (gx.MODE_OZAKI2, 20)
    assert at.read_table(str(bad))[0] == {("emu", "NN", 13, 13, 13): ("ozaki2", 20)}


def test_emu32_entries_pick_the_sgemm_arithmetic(tmp_path, restore_tables):
    """`emu32` lines choose sgemm_'s arithmetic in the automatic mode (native
    or split-BF16x3) per shape bucket, independently of the DGEMM `emu`
    lines; unknown schemes and buckets are skipped, and the Python reader
    and writer agree with the library."""
    path = str(tmp_path / "t.txt")
    e = dict(_entries())
    e[("emu32", "NN", 14, 14, 14)] = "bf16x3"
    e[("emu32", "TN", 13, 13, 13)] = "native"
    e[("emu", "NN", 14, 14, 14)] = ("ozaki2", 19)
    at.write_table(path, e, {}, None)
    got, _, _, _ = at.read_table(path)
    assert got == e
    assert gx.load_tuning(path, "s") == len(e)
    assert gx.tune_lookup_emulated32("N", "N", 16384, 16384, 16384) == gx.MODE_BF16X3
    assert gx.tune_lookup_emulated32("n", "n", 14000, 20000, 16384) == gx.MODE_BF16X3
    assert gx.tune_lookup_emulated32("T", "N", 8192, 8192, 8192) == gx.MODE_NATIVE
    assert gx.tune_lookup_emulated32("T", "T", 16384, 16384, 16384) == gx.MODE_NATIVE
    assert gx.tune_lookup_emulated("N", "N", 16384, 16384, 16384) == (gx.MODE_OZAKI2, 19)
    bad = tmp_path / "bad.txt"
    bad.write_text(f"format 2\nrocm {gx.build_rocm()}\nemu32 NN 14 14 14 fp16\n"
                   "emu32 NN 3 14 14 bf16x3\nemu32 XX 14 14 14 bf16x3\n"
                   "emu32 TT 12 12 12 bf16x3\n")
    assert gx.load_tuning(str(bad), "s") == 1
    assert gx.tune_lookup_emulated32("T", "T", 4096, 4096, 4096) == gx.MODE_BF16X3
    assert gx.tune_lookup_emulated32("N", "N", 16384, 16384, 16384) == gx.MODE_NATIVE
    assert at.read_table(str(bad))[0] == {("emu32", "TT", 12, 12, 12): "bf16x3"}
    # the automatic mode is accepted for sgemm_ too
    assert gx.load("SGEMM.so").sgemm_x_set_mode(gx.MODE_AUTO) == gx.OK
    assert gx.load("SGEMM.so").sgemm_x_set_mode(gx.MODE_NATIVE) == gx.OK


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_strided_entries_pick_the_strided_batch_configuration(tmp_path, dtype, restore_tables):
    """`strided` lines choose the tile configuration of strided batches per
    element type, transpose and shape bucket, but only for batches of at
    least 256 tiles of 128x128; smaller batches, single calls and shapes
    without a line keep the single-problem choice, and a forced
    configuration overrides them.  The Python reader and writer agree with
    the library, which skips malformed lines."""
    path = str(tmp_path / "t.txt")
    want = "f64" if dtype == "d" else "f32"
    e = dict(_entries())
    e[("f64", "NN", 8, 8, 8)] = 3
    e[("f32", "NN", 8, 8, 8)] = 4
    e[("strided", "f64", "NN", 8, 8, 8)] = 5
    e[("strided", "f32", "NN", 8, 8, 8)] = 6
    e[("strided", want, "TN", 10, 10, 10)] = 2
    at.write_table(path, e, {}, None)
    assert at.read_table(path)[0] == e
    assert gx.load_tuning(path, dtype) == len(e)
    single = 3 if dtype == "d" else 4
    strided = 5 if dtype == "d" else 6
    assert gx.tune_lookup(dtype, "N", "N", 256, 256, 256) == single
    # 256^3: four 128x128 tiles a problem, so 64 problems reach 256 tiles
    assert gx.tune_lookup_strided(dtype, "N", "N", 256, 256, 256, 64) == strided
    assert gx.tune_lookup_strided(dtype, "N", "N", 300, 200, 256, 1024) == strided
    assert gx.tune_lookup_strided(dtype, "N", "N", 256, 256, 256, 63) == single
    assert gx.tune_lookup_strided(dtype, "N", "N", 256, 256, 256, 1) == single
    assert gx.tune_lookup_strided(dtype, "T", "N", 1024, 1024, 1024, 4) == 2
    assert gx.tune_lookup_strided(dtype, "T", "N", 1024, 1024, 1024, 3) == \
        gx.tune_lookup(dtype, "T", "N", 1024, 1024, 1024)
    # no line for NT: the single-problem choice
    assert gx.tune_lookup_strided(dtype, "N", "T", 256, 256, 256, 4096) == \
        gx.tune_lookup(dtype, "N", "T", 256, 256, 256)
    assert gx.tune_lookup_strided(dtype, "N", "N", 256, 256, 256, -1) == -1
    gx.set_tile_config(dtype, 7)
    assert gx.tune_lookup_strided(dtype, "N", "N", 256, 256, 256, 4096) == 7
    gx.set_tile_config(dtype, -1)
    bad = tmp_path / "bad.txt"
    bad.write_text(f"format 2\nrocm {gx.build_rocm()}\nstrided {want} NN 8 8 8 {gx.TILE_CONFIGS}\n"
                   f"strided {want} NN 3 8 8 1\nstrided f16 NN 8 8 8 1\nstrided {want} XX 8 8 8 1\n"
                   f"strided {want} NN 9 9 9 8\n")
    assert gx.load_tuning(str(bad), dtype) == 1
    assert gx.tune_lookup_strided(dtype, "N", "N", 512, 512, 512, 4096) == 8
    assert gx.tune_lookup_strided(dtype, "N", "N", 256, 256, 256, 4096) == \
        gx.tune_lookup(dtype, "N", "N", 256, 256, 256)
    assert at.read_table(str(bad))[0] == {("strided", want, "NN", 9, 9, 9): 8}


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_batch_entry_drives_the_batch_configuration(tmp_path, dtype, restore_tables,
                                                    monkeypatch):
    """A `batch <dtype> <cfg>` entry (written and read by the autotuner, read
    by the libraries) gives the grouped-batch kernel configuration; a forced
    configuration and GEMMX_BATCH_CFG take precedence; without an entry the
    default (0) is used."""
    monkeypatch.delenv("GEMMX_BATCH_CFG", raising=False)
    tag = "f64" if dtype == "d" else "f32"
    e = dict(_entries())
    e[("batch", tag)] = 2
    e[("batch", "f32" if dtype == "d" else "f64")] = 3
    path = str(tmp_path / "t.txt")
    at.write_table(path, e, {("batch", tag): "mix"}, None)
    got, notes, _, _ = at.read_table(path)
    assert got == e and notes[("batch", tag)] == "mix"
    try:
        assert gx.load_tuning(path, dtype) == len(e)
        assert gx.batch_config(dtype) == 2
        gx.set_batch_config(dtype, 1)
        assert gx.batch_config(dtype) == 1
        gx.set_batch_config(dtype, -1)
        monkeypatch.setenv("GEMMX_BATCH_CFG", "3")
        assert gx.batch_config(dtype) == 3
        monkeypatch.delenv("GEMMX_BATCH_CFG")
        assert gx.batch_config(dtype) == 2
        at.write_table(path, _entries(), {}, None)
        gx.load_tuning(path, dtype)
        assert gx.batch_config(dtype) == 0
        # out-of-range configurations are skipped by both readers
        with open(path, "a") as f:
            f.write(f"batch {tag} {at.BATCH_CONFIGS}\n")
        assert ("batch", tag) not in at.read_table(path)[0]
        gx.load_tuning(path, dtype)
        assert gx.batch_config(dtype) == 0
    finally:
        gx.set_batch_config(dtype, -1)


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_rocm_release_is_a_checked_key(tmp_path, dtype, restore_tables):
    """A table names the ROCm release whose compiler built the kernels it was
    measured with: the libraries use it only when that is their own release,
    and refuse a table without the key (the previous table stays)."""
    rocm = gx.build_rocm()
    major, minor = (int(v) for v in rocm.split("."))
    good = str(tmp_path / "good.txt")
    at.write_table(good, _entries(), {}, "gfx950")
    assert f"\nrocm {rocm}\n" in open(good).read()
    assert gx.load_tuning(good, dtype) == 5
    want = 6 if dtype == "d" else 2
    body = "f64 NN 13 13 13 1\nf32 NN 13 13 13 1\n"
    for i, head in enumerate((f"format 2\nrocm {major}.{minor + 1}\n",
                              f"format 2\nrocm {major + 1}.{minor}\n",
                              "format 2\n", f"format 2\nrocm {major}\n")):
        bad = tmp_path / f"bad{i}.txt"
        bad.write_text(head + body)
        assert gx.load_tuning(str(bad), dtype) == -1, head
        assert gx.tune_lookup(dtype, "N", "N", 8192, 8192, 8192) == want
    ok = tmp_path / "ok.txt"
    ok.write_text(f"format 2\nrocm {rocm}\n" + body)
    assert gx.load_tuning(str(ok), dtype) == 2
    assert gx.tune_lookup(dtype, "N", "N", 8192, 8192, 8192) == 1
    # the shipped table carries the release the libraries are built with
    shipped = os.path.join(os.path.dirname(at.__file__), "tuning", "gemm_mi355x_tune_gfx950.txt")
    assert at.read_table(shipped)[3] == rocm


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_split_column_forced_split_and_heuristic(tmp_path, dtype, restore_tables):
    """The optional split column of a tile entry sets the number of K ranges
    of a split-K launch; without it a heuristic fills about four workgroups
    per CU when the tiles give fewer than two, with K ranges of at least 256;
    a forced split overrides both; configurations without split-K kernels,
    and shapes with enough tiles, run in one pass."""
    es = 8 if dtype == "d" else 4
    tag = "f64" if dtype == "d" else "f32"
    e = {(tag, "NN", 9, 9, 15): (1, 16), (tag, "NN", 10, 10, 10): (1, 1),
         (tag, "TN", 10, 10, 10): 1, (tag, "NN", 12, 12, 12): (2, 8)}
    path = str(tmp_path / "t.txt")
    at.write_table(path, e, {}, None)
    assert at.read_table(path)[0] == e
    assert gx.load_tuning(path, dtype) == len(e)
    ls = lambda *a: gx.tune_lookup_split(dtype, *a)
    # table: 16 ranges of 4096 at 512 x 512 x 65536 on 64 x 64 tiles
    assert ls("N", "N", 512, 512, 65536) == (16, 16 * 64 * 64 * 64 * es)
    assert ls("N", "N", 1024, 1024, 1024) == (1, 0)  # split 1 in the table
    # no split in the entry: heuristic, 256 tiles of 64x64 -> 4 ranges of 256
    assert ls("T", "N", 1024, 1024, 1024) == (4, 4 * 256 * 64 * 64 * es)
    # configuration 2 has no split-K kernels
    assert ls("N", "N", 4096, 4096, 4096) == (1, 0)
    # outside the table: 8192^2 has enough tiles; 100^2 x 5000: 4 tiles of
    # 64 x 64 -> min(256, 5000 // 256) = 19 ranges of ceil(5000/19) -> 272
    assert ls("N", "N", 8192, 8192, 8192) == (1, 0)
    assert ls("N", "N", 100, 100, 5000)[0] == -(-5000 // 272)
    assert ls("N", "N", 100, 100, 300) == (1, 0)  # K too short to split
    gx.set_split_config(dtype, 4)
    try:
        assert ls("N", "N", 512, 512, 65536)[0] == 4
        assert ls("N", "N", 8192, 8192, 8192)[0] == 4 if gx.tune_lookup(
            dtype, "N", "N", 8192, 8192, 8192) in at.SPLIT_CONFIGS else True
        assert ls("N", "N", 4096, 4096, 4096) == (1, 0)  # configuration 2
        # 100 x 100 x 50: ranges of whole 16-deep slices, so 4 -> 4 ranges of 16
        assert ls("N", "N", 100, 100, 50)[0] == 4
        with pytest.raises(gx.GemmError):
            gx.set_split_config(dtype, 0)
        with pytest.raises(gx.GemmError):
            gx.set_split_config(dtype, at.MAX_SPLIT + 1)
    finally:
        gx.set_split_config(dtype, -1)
    assert ls("N", "N", 512, 512, 65536)[0] == 16


def test_batch_pick_needs_a_clear_lead_over_the_default():
    """The autotuner records a batch configuration other than 0 only when it
    is at least 3% faster and its result agreed with configuration 0's."""
    assert at.batch_pick([2.386, 2.380, 2.363, 2.868], [True] * 4) == 0
    assert at.batch_pick([3.0, 2.8, 3.1, 3.2], [True] * 4) == 1
    assert at.batch_pick([3.0, 2.8, 2.95, 3.2], [True, False, True, True]) == 0
    assert at.batch_pick([3.0, 3.1, 2.5, 3.2], [True, True, True, False]) == 2
