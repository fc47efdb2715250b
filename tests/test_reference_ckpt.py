"""The oracle's checkpoint files (oracle/kukan_oracle_ckpt.c) against the reference's own
rw_wfpot8f.F90 compiled into oracle/_ref/libkukan_ref.so (CPU, one rank).

Files written by the oracle are byte-identical to the reference's; the oracle's readers return
the reference readers' arrays, bit for bit, from reference-written files in every mapping case of
read_wf and read_rho.  The error paths (which abort the reference process) are checked on the
oracle alone, against the reference's messages.  Skipped where the reference library was not
built."""
import numpy as np
import pytest

ref = pytest.importorskip("oracle.ref")
pytestmark = pytest.mark.skipif(not ref.available(), reason="reference library not built")

N = (5, 4, 3)


def _wf(rng, nrc, nums, ncol, neigmx, numk):
    n = int(np.prod(N))
    if nrc:
        v = rng.standard_normal((n, neigmx, nums, numk)) + 1j * rng.standard_normal((n, neigmx, nums, numk))
        return dict(svecre=np.zeros(1), sveccm=np.asfortranarray(v),
                    sval=np.asfortranarray(rng.standard_normal((neigmx, nums + 1 - ncol, numk))))
    v = rng.standard_normal((n, neigmx, nums, numk))
    return dict(svecre=np.asfortranarray(v), sveccm=np.zeros(1, dtype=np.complex128),
                sval=np.asfortranarray(rng.standard_normal((neigmx, nums + 1 - ncol, numk))))


def _dims(nrc, nums, ncol, neigmx, numk):
    return dict(nrc=nrc, nums=nums, ncol=ncol, neigmx=neigmx, numk=numk, ncpx=N[0], ncpy=N[1], ncpz=N[2])


@pytest.mark.parametrize("nrc,nums,numk", [(0, 1, 1), (0, 2, 2), (1, 1, 3), (1, 2, 2)])
def test_write_wf_byte_identical(orc, tmp_path, nrc, nums, numk):
    rng = np.random.default_rng(nrc * 10 + nums + numk)
    a = _wf(rng, nrc, nums, 1, 4, numk)
    files = []
    for f, d in ((orc.call, tmp_path / "o"), (ref.call, tmp_path / "r")):
        d.mkdir()
        f("write_wf", nr1In=3, nr2In=7, chdir=str(d), nrxIn=-1, nryIn=-1, nrzIn=-1, ksym=1,
          **_dims(nrc, nums, 1, 4, numk), **a)
        files.append((d / "wfc.00003.007").read_bytes())
    assert files[0] == files[1]
    assert len(files[0]) > 100


# (file layout) -> (run layout): nrc, nums, ncol, neigmx, numk, ksym
MAPPINGS = [
    ((0, 2, 1, 4, 2, 1), (0, 2, 1, 4, 2, 1)),   # same shape
    ((1, 1, 1, 4, 2, 1), (1, 1, 1, 4, 2, 1)),   # same shape, complex
    ((0, 1, 1, 4, 2, 1), (0, 2, 1, 4, 2, 1)),   # one spin in the file, two in the run
    ((0, 1, 1, 4, 1, 1), (1, 1, 1, 3, 1, 1)),   # real file into a complex run, fewer bands used
    ((1, 2, 1, 3, 1, 1), (1, 2, 2, 6, 1, 1)),   # two collinear spins into a noncollinear run
    ((0, 1, 1, 3, 1, 1), (1, 2, 2, 6, 1, 1)),   # one real spin into a noncollinear run
    ((1, 1, 1, 4, 2, 2), (1, 1, 1, 4, 4, 2)),   # half the k-points, inversion symmetry (ksym 2)
    ((0, 2, 1, 4, 1, 2), (0, 2, 1, 4, 2, 2)),   # the same, real
]


@pytest.mark.parametrize("file,run", MAPPINGS)
def test_read_wf_equals_reference(orc, tmp_path, file, run):
    rng = np.random.default_rng(sum(file) * 7 + sum(run))
    nrc1, nums1, ncol1, neig1, numk1, ksym1 = file
    ref.call("write_wf", nr1In=0, nr2In=0, chdir=str(tmp_path), nrxIn=0, nryIn=0, nrzIn=0, ksym=ksym1,
             **_dims(nrc1, nums1, ncol1, neig1, numk1), **_wf(rng, nrc1, nums1, ncol1, neig1, numk1))
    nrc, nums, ncol, neig, numk, ksym = run
    n = int(np.prod(N))
    outs = []
    for f in (orc.call, ref.call):
        o = dict(svecre=np.full((n, neig, nums, numk) if nrc == 0 else 1, 7.0, order="F"),
                 sveccm=np.full((n, neig, nums, numk) if nrc else 1, 7.0 + 0j, dtype=np.complex128, order="F"),
                 sval=np.full((neig, nums + 1 - ncol, numk), 7.0, order="F"))
        f("read_wf", nr1In=0, nr2In=0, chdir=str(tmp_path), ksym=ksym, key_ksym_inv2=2,
          **_dims(nrc, nums, ncol, neig, numk), **o)
        outs.append(o)
    for k in ("svecre", "sveccm", "sval"):
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=k)


def _rho_args(rng, nspv, nradmx=3, npoint=2, num_atcell=3, natom=4, nmesh=2):
    n = int(np.prod(N))
    sh = (nradmx, npoint, nspv, num_atcell)
    return dict(nspv=nspv, ncpx=N[0], ncpy=N[1], ncpz=N[2], nmesh=nmesh, nradmx=nradmx, npoint=npoint,
                num_atcell=num_atcell, natom=natom, key_natpri_in=2,
                natpri=np.array([2, 1, 2, 0], dtype=np.int32)[:natom],
                rhosmt_i=np.asfortranarray(rng.random((n, nspv))),
                rhotrur_i=np.asfortranarray(rng.random(sh)), rhosmtr_i=np.asfortranarray(rng.random(sh)),
                rhosmt_o=np.asfortranarray(rng.random((n, nspv))),
                rhotrur_o=np.asfortranarray(rng.random(sh)), rhosmtr_o=np.asfortranarray(rng.random(sh)),
                vh_dense=np.asfortranarray(rng.standard_normal(n * nmesh ** 3)))


@pytest.mark.parametrize("nevhist,nrho,nele,npre", [(1, 1, 0, 0), (3, 2, 1, 0), (2, 1, 0, 0), (3, 2, 0, 2)])
def test_write_rhovht_byte_identical(orc, tmp_path, nevhist, nrho, nele, npre):
    rng = np.random.default_rng(nevhist + nrho)
    a = _rho_args(rng, 2)
    listing = []
    for f, d in ((orc.call, tmp_path / "o"), (ref.call, tmp_path / "r")):
        d.mkdir()
        f("write_rhovht", nrIn=2, chdir=str(d), nrho=nrho, nele=nele, npre=npre, nevhist=nevhist,
          nrxIn=-1, nryIn=-1, nrzIn=-1, **a)
        listing.append({p.name: p.read_bytes() for p in sorted(d.iterdir())})
    assert listing[0].keys() == listing[1].keys()
    for k in listing[0]:
        assert listing[0][k] == listing[1][k], k
    if npre != 2:
        assert listing[0]


@pytest.mark.parametrize("nspv_file,nspv_run,nevhist", [(2, 2, 1), (1, 2, 3), (1, 1, -1), (2, 4, 4)])
def test_read_rho_equals_reference(orc, tmp_path, nspv_file, nspv_run, nevhist):
    rng = np.random.default_rng(nspv_file * 3 + nspv_run)
    a = _rho_args(rng, nspv_file)
    ref.call("write_rhovht", nrIn=0, chdir=str(tmp_path), nrho=2, nele=0, npre=0, nevhist=3, nrxIn=0,
             nryIn=0, nrzIn=0, **a)
    n = int(np.prod(N))
    outs = []
    for f in (orc.call, ref.call):
        o = dict(natcell_inf=np.full(3, -1, dtype=np.int32), rhosmt=np.full((n, nspv_run), 7.0, order="F"),
                 rhotrur=np.full((3, 2, nspv_run, 3), 7.0, order="F"),
                 rhosmtr=np.full((3, 2, nspv_run, 3), 7.0, order="F"))
        r = f("read_rho", nrIn=0, chdir=str(tmp_path), nevhist=nevhist, nspv=nspv_run, ncpx=N[0], ncpy=N[1],
              ncpz=N[2], nradmx=3, npoint=2, num_atcell=3, lbrydn=-5, na_atcell=-5, **o)
        o["_out"] = (r["lbrydn"], r["na_atcell"])
        outs.append(o)
    assert outs[0]["_out"] == outs[1]["_out"]
    for k in ("natcell_inf", "rhosmt", "rhotrur", "rhosmtr"):
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=k)


@pytest.mark.parametrize("nevhist,present", [(2, True), (3, True), (1, True), (2, False)])
def test_read_vht_and_filenumbers_equal_reference(orc, tmp_path, nevhist, present):
    rng = np.random.default_rng(5)
    a = _rho_args(rng, 1)
    if present:
        for nr in range(3):
            ref.call("write_rhovht", nrIn=nr, chdir=str(tmp_path), nrho=1, nele=0, npre=0,
                     nevhist=3 if nr < 2 else 2, nrxIn=0, nryIn=0, nrzIn=0, **a)
    outs = []
    for f in (orc.call, ref.call):
        vh = np.full(int(np.prod(N)) * 8, 7.0)
        f("read_vht", nrIn=1, chdir=str(tmp_path), nevhist=nevhist, ncpx=N[0], ncpy=N[1], ncpz=N[2], nmesh=2,
          vh_dense=vh)
        r = f("filenumbers_rhovht", chdir=str(tmp_path), nrrho=-1, nrpot=-1)
        outs.append((vh, r["nrrho"], r["nrpot"]))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    assert outs[0][1:] == outs[1][1:]
    if present:
        assert outs[0][1:] == (2, 3)


def test_reader_errors_carry_reference_messages(orc, tmp_path):
    rng = np.random.default_rng(1)
    ref.call("write_wf", nr1In=0, nr2In=0, chdir=str(tmp_path), nrxIn=0, nryIn=0, nrzIn=0, ksym=1,
             **_dims(0, 1, 1, 3, 1), **_wf(rng, 0, 1, 1, 3, 1))
    n = int(np.prod(N))
    kw = dict(nr1In=0, nr2In=0, chdir=str(tmp_path), ksym=1, key_ksym_inv2=2, svecre=np.zeros(n * 4),
              sveccm=np.zeros(1, dtype=np.complex128), sval=np.zeros(4))
    with pytest.raises(RuntimeError, match="read_wf: unexpected array dimensions in file"):
        orc.call("read_wf", **kw, **dict(_dims(0, 1, 1, 3, 1), ncpx=6))
    with pytest.raises(RuntimeError, match="read_wf: not enough eigenstates in file"):
        orc.call("read_wf", **kw, **_dims(0, 1, 1, 4, 1))
    p = tmp_path / "wfc.00000.000"
    b = bytearray(p.read_bytes())
    good = bytes(b)
    b[4:8] = (4).to_bytes(4, "little")  # version 4
    p.write_bytes(bytes(b))
    with pytest.raises(RuntimeError, match="wrong wfc version"):
        orc.call("read_wf", **kw, **_dims(0, 1, 1, 3, 1))
    p.write_bytes(good[:len(good) - 20])  # a truncated file
    with pytest.raises(RuntimeError, match=r"error found in wfc\.\*"):
        orc.call("read_wf", **kw, **_dims(0, 1, 1, 3, 1))
    with pytest.raises(RuntimeError, match=r"read_rho: nevhist=-1, but no density file rho\.\*"):
        orc.call("read_rho", nrIn=0, chdir=str(tmp_path / "none"), nevhist=-1, nspv=1, ncpx=N[0], ncpy=N[1],
                 ncpz=N[2], nradmx=3, npoint=2, num_atcell=3, natcell_inf=np.zeros(3, dtype=np.int32),
                 rhosmt=np.zeros(n), rhotrur=np.zeros(18), rhosmtr=np.zeros(18))
