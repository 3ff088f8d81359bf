"""Checkpoint files in the reference's format (dmrg_inst.cpp:947-1177).

Fixtures tests/golden/ckpt_su2_ring10/ and ckpt_su2_obc12/ hold the files the
reference's own sweep() wrote (save_mps2, save_ww, save_enr,
dmrg_inst.cpp:940-942); the .npz next to each holds the reference's site
tensors at save time, the tensors its read_mps produced from those files, and
the energy after the resume its driver performs (read_mps, prepare_sweep,
sweep; main/main_dmrg_su2_real.cpp:73-80).

CPU tests pin the oracle's restatement of the format on those files; GPU
tests drive the engine through the C ABI: reading reference files,
byte-identical writing, resuming, and a U(1) / SU(2) round trip of files the
engine's own sweep writes.  The reference's U(1) files have no defined CGC
section (U1Struct::get_nelement leaves ncgc unset), so no U(1) reference file
exists; the U(1) format is pinned by the oracle round trip instead.
"""
import os

import numpy as np
import pytest

import sadmrg_testlib as tl
import sadmrg_oracle as so

HERE = os.path.dirname(os.path.abspath(__file__))
CKPT = {"ckpt_su2_ring10": (1, 10, 16), "ckpt_su2_obc12": (1, 12, 24)}


def _fixture(name):
    xx, yy, D = CKPT[name]
    tens, extra = tl.load_tensors(os.path.join(HERE, "golden", f"{name}.npz"))
    return os.path.join(HERE, "golden", name), xx * yy, D, tens, extra


def _site_files(d, ns, D, i):
    base = os.path.join(d, f"uu-{ns}-{D}-0-0-{i}")
    with open(base + ".dat") as f:
        dat = f.read()
    with open(base + ".bin", "rb") as f:
        binary = f.read()
    return dat, binary


def _same_blocks(t, ref):
    """Block dictionaries equal element by element (absent = zero)."""
    assert t.bonds == ref.bonds and t.locspin == ref.locspin
    for b in set(t.blocks) | set(ref.blocks):
        x = t.blocks.get(b)
        y = ref.blocks.get(b)
        if x is None:
            x = np.zeros_like(y)
        if y is None:
            y = np.zeros_like(x)
        assert np.array_equal(x, y), b


# ---- oracle vs reference-written files (CPU) ---------------------------------------
@pytest.mark.parametrize("name", list(CKPT))
def test_oracle_reads_reference_files(name):
    """read_mps restated: every site equals the tensor the reference's own
    read_mps built from the same files, and the CGC section holds the
    canonical fill_block tensors of the blocks."""
    d, ns, D, tens, extra = _fixture(name)
    assert int(extra["read_pass"]) == 1
    for i in range(ns):
        dat, binary = _site_files(d, ns, D, i)
        t, cgc = so.read_mps_site(dat, binary, 1, 0)
        _same_blocks(t, so.from_packed(tens[f"read_uu_{i}"]))
        _same_blocks(t, so.from_packed(tens[f"uu_{i}"]))
        canon = np.concatenate([so.cgc_tensor(*t.labels(b)).ravel(order="F")
                                for b in range(t.nten()) if t.allowed(b)])
        assert np.array_equal(cgc, canon)


@pytest.mark.parametrize("name", list(CKPT))
def test_oracle_writes_reference_bytes(name):
    """save_mps2 restated: the .dat text and .bin bytes of every site equal
    the reference's files; save_enr writes %.16g lines; save_ww writes
    bondd * phdim doubles of ww (phdim = 2 for spin-1/2)."""
    d, ns, D, tens, extra = _fixture(name)
    for i in range(ns):
        dat, binary = _site_files(d, ns, D, i)
        txt, by = so.mps_site_files(so.from_packed(tens[f"uu_{i}"]))
        assert txt == dat
        assert by == binary
    with open(os.path.join(d, f"enr-{ns}-{D}-0-0.dat")) as f:
        assert f.read() == "%.16g\n" % float(extra["e_saved"])
    ww = np.fromfile(os.path.join(d, f"ww-{ns}-{D}-0-0.bin"))
    assert len(ww) == 2 * D
    assert np.array_equal(ww, np.asarray(extra["ww"])[:2 * D])


def test_oracle_u1_roundtrip():
    """U(1) site tensors (reference mid-sweep state of a 10-site ring) written
    and read back by the restated format are unchanged; the file ends after
    the data section."""
    tens, _ = tl.load_tensors(os.path.join(HERE, "golden", "dmrg_u1_ring10.npz"))
    names = [k for k in tens if k.startswith("S_uu_")]
    assert names
    for k in names:
        t = so.from_packed(tens[k])
        if t.nbond != 3:
            continue
        txt, by = so.mps_site_files(t)
        back, cgc = so.read_mps_site(txt, by, 0, 0)
        assert cgc.size == 0
        nd = sum(int(np.prod(t.dims(b))) for b in range(t.nten()) if t.allowed(b))
        assert len(by) == 8 * nd
        present = {b: v for b, v in t.blocks.items()}
        for b, v in back.blocks.items():
            assert np.array_equal(v, present.get(b, np.zeros_like(v)))


# ---- engine (GPU) --------------------------------------------------------------------
def _engine(sad, ctx, extra, sym=1):
    _, _, xx, yy, D, ns = (int(x) for x in extra["params"])
    d = sad.DMRG(ctx, sym, 0, xx, yy, D)
    d.set_couplings(extra["hmap"], extra["coup"])
    d.set_max_dcut(D)
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CKPT))
def test_engine_reads_reference_files(ctx, sad, name):
    """read_mps(read1, read2): every site equals the reference's read result
    bit for bit; the energy is the file's last line; ww the file's first
    read1 values."""
    d_, ns, D, tens, extra = _fixture(name)
    d = _engine(sad, ctx, extra)
    assert d.read_mps(d_, D, 0)
    for i in range(ns):
        got = so.from_packed(tl.from_dev(d.tensor("uu", i)))
        _same_blocks(got, so.from_packed(tens[f"read_uu_{i}"]))
    assert d.energy() == float(extra["e_read"])
    ww = np.fromfile(os.path.join(d_, f"ww-{ns}-{D}-0-0.bin"))
    assert np.array_equal(d.ww(), ww[:D])


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CKPT))
def test_engine_writes_reference_bytes(ctx, sad, name, tmp_path):
    """After reading the reference's files, save_mps2 / save_ww / save_enr
    reproduce them: site files byte for byte, the first read1 values of ww
    (what read_ww keeps), the energy line."""
    d_, ns, D, tens, extra = _fixture(name)
    d = _engine(sad, ctx, extra)
    assert d.read_mps(d_, D, 0)
    out = str(tmp_path)
    d.save_mps(out, 0)
    d.save_ww(out)
    d.save_enr(out)
    for i in range(ns):
        for ext in (".dat", ".bin"):
            with open(os.path.join(d_, f"uu-{ns}-{D}-0-0-{i}{ext}"), "rb") as f:
                ref = f.read()
            with open(os.path.join(out, f"uu-{ns}-{D}-0-0-{i}{ext}"), "rb") as f:
                got = f.read()
            assert got == ref, (i, ext)
    ww_ref = np.fromfile(os.path.join(d_, f"ww-{ns}-{D}-0-0.bin"))
    ww = np.fromfile(os.path.join(out, f"ww-{ns}-{D}-0-0.bin"))
    assert len(ww) == len(ww_ref) and np.array_equal(ww[:D], ww_ref[:D])
    with open(os.path.join(out, f"enr-{ns}-{D}-0-0.dat")) as f, \
            open(os.path.join(d_, f"enr-{ns}-{D}-0-0.dat")) as g:
        assert f.read() == g.read()


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CKPT))
def test_engine_resume_matches_reference(ctx, sad, name):
    """The reference driver's resume (read_mps, prepare_sweep, sweep) from the
    reference's files: same energy as the reference's resumed sweep."""
    d_, ns, D, tens, extra = _fixture(name)
    d = _engine(sad, ctx, extra)
    assert d.read_mps(d_, D, 0)
    d.prepare_sweep()
    d.sweep()
    e, e_ref = d.energy(), float(extra["e_resume"])
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref), (e, e_ref)


@pytest.mark.gpu
@pytest.mark.parametrize("sym,D", [(1, 16), (0, 24)])
def test_sweep_checkpoint_roundtrip(ctx, sad, sym, D, tmp_path):
    """sweep() with a checkpoint directory writes the reference's files
    (dmrg_inst.cpp:940-942); the oracle reads every site back equal to the
    engine's tensor, and a fresh engine resumes from them to the same
    energy as the original continuing."""
    out = str(tmp_path)
    sad.lib().sadmrg_rng_seed(42)
    d = sad.DMRG(ctx, sym, 0, 1, 10, D)
    d.set_max_dcut(D)
    d.idmrg()
    d.set_checkpoint_dir(out)
    d.sweep()
    d.set_checkpoint_dir(None)
    ns = 10
    for i in range(ns):
        dat, binary = _site_files(out, ns, D, i)
        t, _ = so.read_mps_site(dat, binary, sym, 0)
        _same_blocks(t, so.from_packed(tl.from_dev(d.tensor("uu", i))))
    with open(os.path.join(out, f"enr-{ns}-{D}-0-0.dat")) as f:
        assert abs(float(f.read().split()[-1]) - d.energy()) <= 1e-15 * abs(d.energy())
    r = sad.DMRG(ctx, sym, 0, 1, 10, D)
    r.set_max_dcut(D)
    assert r.read_mps(out, D, 0)
    assert abs(r.energy() - d.energy()) <= 1e-15 * abs(d.energy())  # 16 digits in the file
    r.prepare_sweep()
    d.prepare_sweep()
    r.sweep()
    d.sweep()
    assert abs(r.energy() - d.energy()) <= 1e-12 * abs(d.energy())


@pytest.mark.gpu
def test_checkpoint_errors(ctx, sad, tmp_path):
    """Missing files clear the pass flag (as the reference's read_mps);
    a short .bin or a non-canonical CGC block is an error (the reference
    reads such files unchecked)."""
    d_, ns, D, tens, extra = _fixture("ckpt_su2_ring10")
    d = _engine(sad, ctx, extra)
    assert not d.read_mps(str(tmp_path), D, 0)
    import shutil
    bad = tmp_path / "bad"
    shutil.copytree(d_, bad)
    site = bad / f"uu-{ns}-{D}-0-0-4.bin"
    raw = site.read_bytes()
    site.write_bytes(raw[:8])  # one element of a data section that holds several
    with pytest.raises(sad.SadmrgError, match="shorter"):
        d.read_mps(str(bad), D, 0)
    arr = np.frombuffer(raw, np.float64).copy()
    arr[-1] += 0.5  # last CGC element
    site.write_bytes(arr.tobytes())
    with pytest.raises(sad.SadmrgError, match="CGC"):
        d.read_mps(str(bad), D, 0)
