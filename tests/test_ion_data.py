"""The data-file network builders on the CPU (dengo_amd.ion_by_ion):

* the reference's own HDF5 data files are read through the HIP library's
  HDF5 access (dengo_h5_*, no h5py): a committed copy of
  input/cooling/o_ion_by_ion_cooling.h5 and, when the reference tree is
  present, every oxygen input file, equal the .npz fixtures written from them;
* Gnat & Ferland cooling tables: np.interp inside the data, Gaussian fall-off
  to tiny at the ends of the grid (dengo/reaction_classes.py:392-432);
* HM12 photo tables: np.interp inside the data, straight lines in
  log10 z - log10 rate beyond it (:215-246);
* the O network built from them, generated, equals the oracle's mass-action
  ion chain with the photo rates off, and numpy on its printed expressions."""
import os

import numpy as np
import pytest

import oracle
from dengo_amd import codegen, ion_by_ion, tables
from dengo_amd.ics import oxygen_ics

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "tests", "golden", "ion_data")
REF_H5 = os.path.join(ROOT, "tests", "golden", "reference_h5", "o_ion_by_ion_cooling.h5")
REF = "/root/reference"
needs_h5 = pytest.mark.skipif(not tables.hdf5_available(), reason="libhdf5 not loadable")


@pytest.fixture(scope="module")
def oxynet():
    from dengo_amd.networks import oxygen_data_network
    with np.errstate(all="ignore"):
        net = oxygen_data_network(DATA)
        return net, net.table_block()


@needs_h5
def test_reference_hdf5_file_reads_through_dengo_h5():
    names = ["T"] + ["o_%d" % k for k in range(1, 10)]
    got, _ = tables.read_h5(REF_H5, names)
    fix = np.load(os.path.join(DATA, "input", "cooling", "o_ion_by_ion_cooling.npz"))
    for k in names:
        assert np.array_equal(got[k], fix[k]), k
    assert got["T"].size == 176 and got["T"][0] == 100.0
    assert np.all(np.diff(got["T"]) > 0)


@needs_h5
@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "input")), reason="reference tree absent")
def test_fixtures_equal_the_reference_input_files():
    for kind, base in (("cooling", "o_ion_by_ion_cooling"),
                       ("photoionization", "o_ion_by_ion_photoionization_HM12"),
                       ("photoheating", "o_ion_by_ion_photoheating_HM12")):
        fix = np.load(os.path.join(DATA, "input", kind, base + ".npz"))
        got = ion_by_ion.read_data(os.path.join(REF, "input", kind, base), list(fix.files))
        for k in fix.files:
            assert np.array_equal(got[k], fix[k]), (kind, k)


def test_cooling_tables_interpolate_and_fall_off(oxynet):
    net, tab = oxynet
    names = net.table_names()
    fix = np.load(os.path.join(DATA, "input", "cooling", "o_ion_by_ion_cooling.npz"))
    T, dT = net.T, fix["T"]
    inside = (T >= dT[0]) & (T <= dT[-1])
    for k in range(1, 10):
        row = tab[names.index("O_%d_c_O_%d_c" % (k, k))]
        S = fix["o_%d" % k]
        assert np.array_equal(row[inside], np.interp(T[inside], dT, S))
        above = T > dT[-1]
        assert above.sum() > 5
        # Gaussian fall-off from the last data point to tiny (1e-30) at T[-1]
        assert np.isclose(row[-1], 1e-30, rtol=1e-9, atol=0)
        assert np.all(np.diff(row[above]) < 0)
        x = (T[above] - dT[-1]) / (T[-1] - dT[-1])
        assert np.allclose(np.log(row[above] / S[-1]), x ** 2 * np.log(1e-30 / S[-1]),
                           rtol=1e-9, atol=1e-9)


def test_photo_tables_interpolate_and_extrapolate(oxynet):
    net, tab = oxynet
    names = net.table_names()
    z, nz = net.z, len(net.z)
    for kind, fmt in (("photoionization", "O_%d_pi"), ("photoheating", "O_%d_ph_O_%d_ph")):
        fix = np.load(os.path.join(DATA, "input", kind, "o_ion_by_ion_%s_HM12.npz" % kind))
        dz = fix["z"]
        for k in range(1, 9):
            row = tab[names.index(fmt % ((k,) if kind == "photoionization" else (k, k)))][:nz]
            S = fix["o_%d" % k]
            inside = (z >= dz[0]) & (z <= dz[-1])
            assert np.allclose(row[inside], np.interp(z[inside], dz, S), rtol=1e-13, atol=0)
            above = z > dz[-1]
            assert above.sum() >= 1
            slope = (np.log10(S[-1]) - np.log10(S[-2])) / (np.log10(dz[-1]) - np.log10(dz[-2]))
            want = np.log10(S[-1]) + (np.log10(z[above]) - np.log10(dz[-1])) * slope
            assert np.allclose(np.log10(row[above]), want, rtol=1e-12, atol=0)


def test_data_network_is_generated_and_matches_the_ion_chain_oracle(oxynet):
    net, tab = oxynet
    assert codegen.device_network_kind(net) == "gen"
    g = oracle.gen_net("oxydata")
    _, n, _ = oxygen_ics(400, seed=71)
    TB = net.T_bounds
    # photo rates off (z outside 0 <= z < 9): the mass-action ion chain with
    # the data cooling tables
    names = net.table_names()
    mnet = oracle.ion_chain_net(8, 15.9994)
    chain = []
    for k in range(1, 10):
        if k < 9:
            chain.append("O_%d_i" % k)
        if k > 1:
            chain.append("O_%d_r" % k)
    chain.sort(key=str.lower)
    chain += ["O_%d_c_O_%d_c" % (k, k) for k in range(1, 10)]
    mtab = np.stack([tab[names.index(t)] for t in chain])
    f1, J1, T1, r1, _ = oracle.gen_eval(g, n, tab, TB, z=-1.0)
    f0, J0, T0, r0, _ = oracle.ma_eval(mnet, n, mtab, TB)
    assert np.array_equal(T0, T1)
    scale = np.maximum(np.abs(f0).max(axis=1, keepdims=True), 1e-300)
    assert np.all(np.abs(f1[:, :10] - f0[:, :10]) <= 1e-12 * scale[:, :1])
    assert np.allclose(f1[:, 10], f0[:, 10], rtol=1e-11, atol=0)
    Js = np.maximum(np.abs(J0).max(axis=(1, 2), keepdims=True), 1e-300)
    assert np.all(np.abs(J1 - J0) <= 1e-12 * Js)
    # with the background on, O_k loses Gamma_k n_k to O_{k+1}
    f2, _, _, r2, _ = oracle.gen_eval(g, n, tab, TB, z=3.0)
    gam = r2[:, names.index("O_1_pi")]
    assert np.all(gam > 0)
    d = f2[:, 0] - f1[:, 0]
    assert np.allclose(d, -gam * n[:, 0], rtol=1e-6, atol=1e-12 * scale[:, 0])
