"""Checkpoint files through the GPU library: kukan_write_wf / write_rhovht produce the same bytes
as the oracle (which the reference's own writers pin, tests/test_reference_ckpt.py), from host
arrays and from arrays in HBM; the readers, into HBM, return the oracle readers' arrays from
files the reference itself wrote (tests/golden/ckpt, made by tests/golden/make_ckpt_golden.py)
in each mapping case; and a wave-function set of several hundred MB round-trips through a file
bit for bit, with records larger than the staging buffers."""
import os

import numpy as np
import pytest

from tests.golden.make_ckpt_golden import N, RHO, WF_CASES, case_data

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ckpt")


@pytest.fixture(scope="module")
def kk():
    from rspacedist_amd import kukan
    kukan.init(0)
    yield kukan


def _dims(nrc, nums, ncol, neigmx, numk):
    return dict(nrc=nrc, nums=nums, ncol=ncol, neigmx=neigmx, numk=numk, ncpx=N[0], ncpy=N[1], ncpz=N[2])


@pytest.mark.parametrize("name", sorted(WF_CASES))
@pytest.mark.parametrize("on_device", [False, True])
def test_write_wf_same_bytes_as_oracle(kk, orc, tmp_path, name, on_device):
    nrc, nums, ncol, neig, numk, ksym = WF_CASES[name]
    a = case_data(name)
    if on_device:
        a = {k: kk.DeviceArray.from_numpy(v) for k, v in a.items()}
    (tmp_path / "g").mkdir()
    (tmp_path / "o").mkdir()
    kk.call("write_wf", nr1In=4, nr2In=1, chdir=str(tmp_path / "g"), nrxIn=-1, nryIn=-1, nrzIn=-1, ksym=ksym,
            **_dims(nrc, nums, ncol, neig, numk), **a)
    orc.call("write_wf", nr1In=4, nr2In=1, chdir=str(tmp_path / "o"), nrxIn=-1, nryIn=-1, nrzIn=-1, ksym=ksym,
             **_dims(nrc, nums, ncol, neig, numk), **case_data(name))
    g = (tmp_path / "g" / "wfc.00004.001").read_bytes()
    assert g == (tmp_path / "o" / "wfc.00004.001").read_bytes()
    # and the same bytes as the reference's own file of these data
    assert g == open(os.path.join(GOLD, name, "wfc.00000.000"), "rb").read()


# file case -> run layout (nrc, nums, ncol, neigmx, numk, ksym)
READS = [("wf_real_2spin", (0, 2, 1, 4, 1, 1)), ("wf_cplx_halfk", (1, 1, 1, 3, 4, 2)),
         ("wf_real_1spin", (1, 2, 1, 3, 1, 1)), ("wf_cplx_2spin", (1, 2, 2, 6, 1, 1)),
         ("wf_real_2spin", (1, 2, 1, 4, 1, 1))]


@pytest.mark.parametrize("name,run", READS)
@pytest.mark.parametrize("on_device", [False, True])
def test_read_wf_reference_files(kk, orc, name, run, on_device):
    nrc, nums, ncol, neig, numk, ksym = run
    n = int(np.prod(N))
    out = []
    for f in (kk.call, orc.call):
        o = dict(svecre=np.full((n, neig, nums, numk) if nrc == 0 else 1, 7.0, order="F"),
                 sveccm=np.full((n, neig, nums, numk) if nrc else 1, 7.0 + 0j, dtype=np.complex128, order="F"),
                 sval=np.full((neig, nums + 1 - ncol, numk), 7.0, order="F"))
        if on_device and f is kk.call:
            d = {k: kk.DeviceArray.from_numpy(v) for k, v in o.items()}
            f("read_wf", nr1In=0, nr2In=0, chdir=os.path.join(GOLD, name), ksym=ksym, key_ksym_inv2=2,
              **_dims(nrc, nums, ncol, neig, numk), **d)
            o = {k: v.numpy() for k, v in d.items()}
        else:
            f("read_wf", nr1In=0, nr2In=0, chdir=os.path.join(GOLD, name), ksym=ksym, key_ksym_inv2=2,
              **_dims(nrc, nums, ncol, neig, numk), **o)
        out.append(o)
    for k in ("svecre", "sveccm", "sval"):
        np.testing.assert_array_equal(out[0][k], out[1][k], err_msg=k)
    if run[:5] == WF_CASES[name][:5]:  # same layout: the data that were written
        src = case_data(name)
        np.testing.assert_array_equal(out[0]["sveccm" if nrc else "svecre"], src["sveccm" if nrc else "svecre"])


@pytest.mark.parametrize("on_device", [False, True])
def test_rho_vht_files(kk, orc, tmp_path, on_device):
    n = int(np.prod(N))
    a = case_data("rhovht")
    dims = dict(nspv=RHO["nspv"], ncpx=N[0], ncpy=N[1], ncpz=N[2], nmesh=RHO["nmesh"], nradmx=RHO["nradmx"],
                npoint=RHO["npoint"], num_atcell=RHO["num_atcell"], natom=RHO["natom"], key_natpri_in=2)
    src = {k: (kk.DeviceArray.from_numpy(v) if on_device and k != "natpri" else v) for k, v in a.items()}
    kk.call("write_rhovht", nrIn=0, chdir=str(tmp_path), nrho=2, nele=1, npre=0, nevhist=3, nrxIn=0, nryIn=0,
            nrzIn=0, **dims, **src)
    for f in ("rho.00000", "r_o.00000", "vht.00000"):
        assert (tmp_path / f).read_bytes() == open(os.path.join(GOLD, "rhovht", f), "rb").read(), f
    sh = (RHO["nradmx"], RHO["npoint"], RHO["nspv"], RHO["num_atcell"])
    for nevhist in (1, 4):
        o = dict(natcell_inf=np.zeros(2, dtype=np.int32), rhosmt=np.zeros((n, 2), order="F"),
                 rhotrur=np.zeros(sh, order="F"), rhosmtr=np.zeros(sh, order="F"))
        d = {k: (kk.DeviceArray.from_numpy(v) if on_device and k != "natcell_inf" else v) for k, v in o.items()}
        r = kk.call("read_rho", nrIn=0, chdir=os.path.join(GOLD, "rhovht"), nevhist=nevhist, nspv=2, ncpx=N[0],
                    ncpy=N[1], ncpz=N[2], nradmx=RHO["nradmx"], npoint=RHO["npoint"], num_atcell=2, **d)
        got = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in d.items()}
        assert (r["lbrydn"], r["na_atcell"]) == (1, 2)
        sfx = "_i" if nevhist == 1 else "_o"
        np.testing.assert_array_equal(got["rhosmt"], a["rhosmt" + sfx])
        np.testing.assert_array_equal(got["rhotrur"], a["rhotrur" + sfx])
        np.testing.assert_array_equal(got["natcell_inf"], [1, 3])
    vh = np.zeros(n * 8)
    dv = kk.DeviceArray.from_numpy(vh) if on_device else vh
    kk.call("read_vht", nrIn=0, chdir=os.path.join(GOLD, "rhovht"), nevhist=2, ncpx=N[0], ncpy=N[1], ncpz=N[2],
            nmesh=2, vh_dense=dv)
    np.testing.assert_array_equal(dv.numpy() if on_device else dv, a["vh_dense"])
    r = kk.call("filenumbers_rhovht", chdir=str(tmp_path))
    assert (r["nrrho"], r["nrpot"]) == (1, 1)


def test_errors_carry_reference_messages(kk, tmp_path):
    n = int(np.prod(N))
    kw = dict(nr1In=0, nr2In=0, ksym=1, key_ksym_inv2=2, svecre=np.zeros(n * 8),
              sveccm=np.zeros(1, dtype=np.complex128), sval=np.zeros(8))
    with pytest.raises(kk.KukanError, match="read_wf: unexpected array dimensions in file"):
        kk.call("read_wf", chdir=os.path.join(GOLD, "wf_real_2spin"), **kw, **_dims(0, 2, 1, 4, 2))
    with pytest.raises(kk.KukanError, match="read_wf: not enough eigenstates in file"):
        kk.call("read_wf", chdir=os.path.join(GOLD, "wf_real_1spin"), **kw, **_dims(0, 1, 1, 5, 1))
    with pytest.raises(kk.KukanError, match=r"cannot open wfc\.\* file"):
        kk.call("read_wf", chdir=str(tmp_path / "none"), **kw, **_dims(0, 1, 1, 4, 1))
    b = bytearray(open(os.path.join(GOLD, "wf_real_1spin", "wfc.00000.000"), "rb").read())
    (tmp_path / "t").mkdir()
    (tmp_path / "t" / "wfc.00000.000").write_bytes(bytes(b[:-30]))  # truncated
    with pytest.raises(kk.KukanError, match=r"error found in wfc\.\*"):
        kk.call("read_wf", chdir=str(tmp_path / "t"), **kw, **_dims(0, 1, 1, 4, 1))


def test_device_round_trip_streaming(kk, tmp_path):
    """64^3 complex x 24 bands x 2 k-points (about 200 MB) from HBM to a file and back, and a
    dense potential of 200^3 points whose one record (64 MB) is larger than a staging buffer."""
    n3, neig, numk = 64 ** 3, 24, 2
    src = kk.DeviceArray((n3, neig, 1, numk), "complex128").fill_normal(11)
    sval = kk.DeviceArray((neig, 1, numk)).fill_normal(12)
    kw = dict(nrc=1, nums=1, ncol=1, neigmx=neig, numk=numk, ncpx=64, ncpy=64, ncpz=64)
    kk.call("write_wf", nr1In=0, nr2In=0, chdir=str(tmp_path), nrxIn=0, nryIn=0, nrzIn=0, ksym=1,
            svecre=None, sveccm=src, sval=sval, **kw)
    assert os.path.getsize(tmp_path / "wfc.00000.000") == 12 + 40 + 24 + 12 + 8 + 8 * neig * numk + \
        neig * numk * (8 + 16 * n3)
    dst = kk.DeviceArray((n3, neig, 1, numk), "complex128")
    sv2 = kk.DeviceArray((neig, 1, numk))
    kk.call("read_wf", nr1In=0, nr2In=0, chdir=str(tmp_path), ksym=1, key_ksym_inv2=2, svecre=None,
            sveccm=dst, sval=sv2, **kw)
    a, b = src.numpy(), dst.numpy()
    assert np.array_equal(a.ravel(order="F").view(np.uint64), b.ravel(order="F").view(np.uint64))
    assert np.array_equal(sval.numpy(), sv2.numpy())
    vh = kk.DeviceArray((200 ** 3,)).fill_normal(13)
    z = np.zeros(1)
    kk.call("write_rhovht", nrIn=0, chdir=str(tmp_path), nrho=1, nele=0, npre=0, nevhist=2, nspv=1, ncpx=100,
            ncpy=100, ncpz=100, nmesh=2, nradmx=1, npoint=1, num_atcell=1, nrxIn=0, nryIn=0, nrzIn=0, natom=1,
            key_natpri_in=2, natpri=np.zeros(1, dtype=np.int32), rhosmt_i=z, rhotrur_i=z, rhosmtr_i=z,
            rhosmt_o=z, rhotrur_o=z, rhosmtr_o=z, vh_dense=vh)
    back = kk.DeviceArray((200 ** 3,))
    kk.call("read_vht", nrIn=0, chdir=str(tmp_path), nevhist=2, ncpx=100, ncpy=100, ncpz=100, nmesh=2,
            vh_dense=back)
    assert np.array_equal(vh.numpy(), back.numpy())
