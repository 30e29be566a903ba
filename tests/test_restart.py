"""eT restart files on the host (et_amd.restart), no GPU.

* eT's own v1.0 restart directories (tests/restart_ccsd_v1.0, restart_ccsd_from_cc2_v1.0),
  rebuilt byte for byte from the fixtures, convert to the current amplitude-file
  format: zero energy, int64 block lengths, the records' doubles; excited states
  carry eT's excitation energies;
  the CC2 t file (singles only) reads back as a partial vector;
* amplitude files round-trip and reject blocks of the wrong length as eT's
  read_ does; orbital files round-trip and check their dimensions;
* the packed (contra/co)variant maps equal 2 X_aibj - X_biaj and (2 X_aibj + X_biaj) / 3
  formed on the full 4-index array, and are inverse to each other, so a right
  vector restarted as left and back is unchanged.
"""
import os

import numpy as np
import pytest

import v1_0_files as V
from et_amd import restart as R
from oracle import ccsd as O


def test_rebuilt_v1_0_directories_are_ets_files(tmp_path):
    for which in ('ccsd', 'cc2'):
        d = V.build(str(tmp_path / which), which)
        assert sorted(os.listdir(d))            # build() raises on any size / hash mismatch


def test_v1_0_ccsd_directory_converts_to_amplitude_files(tmp_path):
    old = V.build(str(tmp_path / 'old'))
    new = str(tmp_path / 'new')
    R.convert_v1_0_directory(old, new)
    v = V.vectors()
    n_t2 = V.N_T1 * (V.N_T1 + 1) // 2
    raw = open(os.path.join(new, 't'), 'rb').read()
    assert len(raw) == 8 + 8 + 8 * V.N_T1 + 8 + 8 * n_t2
    assert raw[:8] == bytes(8)
    assert int.from_bytes(raw[8:16], 'little') == V.N_T1
    assert int.from_bytes(raw[16 + 8 * V.N_T1:24 + 8 * V.N_T1], 'little') == n_t2
    files = R.CCSDFiles(new, V.N_O, V.N_V)
    for name, f in (('t', files.t), ('tbar', files.tbar)):
        x, e, n_read = f.read()
        assert n_read == V.N_T1 + n_t2 and e == 0.0
        assert np.array_equal(x, v[name])
    for k in (1, 2, 3):
        for side in ('right', 'left'):
            x, e, n_read = files.side(side, k).read()
            assert n_read == V.N_T1 + n_t2
            assert np.array_equal(x, v[side[0] + str(k)])
            assert e == v['excitation_energies'][k - 1]
    C, eps = files.orbitals.read(n_ao=29, n_mo=29)
    assert np.array_equal(eps, v['orbital_energies'])
    raw_C = bytes.fromhex(V.fixture()['sets']['ccsd']['orbital_coefficients']['hex'])[4:-4]
    assert np.array_equal(C, np.frombuffer(raw_C, '<f8').reshape(29, 29).T)
    with pytest.raises(R.RestartError, match='Number of AOs'):
        files.orbitals.read(n_ao=30)


def test_v1_0_cc2_ground_state_is_singles_only(tmp_path):
    """CC2 v1.0 ground-state files hold singles only: CCSD reads n_t1 amplitudes."""
    old = V.build(str(tmp_path / 'old'), 'cc2')
    new = str(tmp_path / 'new')
    R.convert_v1_0_directory(old, new)
    assert os.path.getsize(os.path.join(new, 't')) == 16 + 8 * V.N_T1
    x, e, n_read = R.CCSDFiles(new, V.N_O, V.N_V).t.read()
    assert n_read == V.N_T1
    t1 = np.frombuffer(bytes.fromhex(V.fixture()['sets']['cc2']['t']['hex'])[4:-4], '<f8')
    assert np.array_equal(x[:V.N_T1], t1) and not x[V.N_T1:].any()


def test_converter_rejects_wrong_sizes(tmp_path):
    old = V.build(str(tmp_path / 'old'))
    with open(os.path.join(old, 'r_002'), 'ab') as f:
        f.write(bytes(8))
    with pytest.raises(R.RestartError, match='dimensions in cc_restart_file'):
        R.convert_v1_0_directory(old, str(tmp_path / 'new'))
    with open(os.path.join(old, 'r_002'), 'r+b') as f:          # a record cut short
        f.truncate(os.path.getsize(os.path.join(old, 'r_002')) - 12)
    with pytest.raises(R.RestartError, match='not a Fortran sequential file'):
        R.convert_v1_0_directory(old, str(tmp_path / 'new'))
    with pytest.raises(R.RestartError, match='would overwrite'):
        R.convert_v1_0_directory(old, old)


def test_amplitude_file_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(0)
    f = R.AmplitudeFile(str(tmp_path / 'x'), [4, 10])
    assert not f.exists()
    x = rng.normal(size=14)
    f.save(x, -1.25)
    y, e, n = f.read()
    assert np.array_equal(x, y) and e == -1.25 and n == 14
    assert os.path.getsize(f.path) == 8 + 8 + 32 + 8 + 80
    # a singles-only file (eT's CCS t): reading stops at the missing block header
    R.AmplitudeFile(f.path, [4]).save(x[:4], 0.5)
    y, e, n = f.read()
    assert n == 4 and np.array_equal(y[:4], x[:4]) and not y[4:].any()
    # a block of another length
    R.AmplitudeFile(f.path, [4, 9]).save(x[:13])
    with pytest.raises(R.RestartError, match='Expected 10, found 9'):
        f.read()
    # truncated inside a block
    with open(f.path, 'wb') as g:
        g.write(np.float64(0).tobytes() + np.int64(4).tobytes() + bytes(16))
    with pytest.raises(R.RestartError, match='ends inside'):
        f.read()
    with pytest.raises(R.RestartError):
        f.save(x[:5])


def test_orbital_file_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    C, eps = rng.normal(size=(7, 5)), np.sort(rng.normal(size=5))
    f = R.OrbitalFile(str(tmp_path / 'orbital_coefficients'))
    f.save(C, eps)
    raw = open(f.path, 'rb').read()
    assert np.frombuffer(raw[:16], np.int64).tolist() == [7, 5]
    assert np.array_equal(np.frombuffer(raw[16 + 40:], np.float64), C.ravel(order='F'))
    C2, eps2 = f.read(7, 5)
    assert np.array_equal(C, C2) and np.array_equal(eps, eps2)
    with pytest.raises(R.RestartError, match='Number of MOs'):
        f.read(7, 6)


@pytest.mark.parametrize('n_o,n_v', [(1, 1), (2, 3), (3, 2), (4, 5)])
def test_packed_variance_maps_against_four_index_arrays(n_o, n_v):
    rng = np.random.default_rng(n_o * 10 + n_v)
    n1 = n_o * n_v
    x2 = rng.normal(size=n1 * (n1 + 1) // 2)
    X = O.unpack_t2(x2, n_v, n_o)                                     # X[a, i, b, j]
    swapped = X.transpose(2, 1, 0, 3)                                 # X[b, i, a, j]
    assert np.allclose(R.packed_contravariant(x2, n_v, n_o), O.pack_aibj(2 * X - swapped), atol=1e-14)
    assert np.allclose(R.packed_covariant(x2, n_v, n_o), O.pack_aibj((2 * X + swapped) / 3), atol=1e-14)
    assert np.allclose(R.packed_covariant(R.packed_contravariant(x2, n_v, n_o), n_v, n_o), x2, atol=1e-13)


def test_restart_vector_side_changes():
    n_o, n_v = 3, 4
    n1 = n_o * n_v
    n = n1 + n1 * (n1 + 1) // 2
    rng = np.random.default_rng(5)
    r = rng.normal(size=n)
    left = R.restart_vector(r, n, n_o, n_v, 'right', 'left')
    assert np.allclose(left[:n1], 2 * r[:n1])
    assert np.allclose(left[n1:], 2 * R.packed_contravariant(r[n1:], n_v, n_o))
    back = R.restart_vector(left, n, n_o, n_v, 'left', 'right')
    assert np.allclose(back, r, atol=1e-13)
    assert np.array_equal(R.restart_vector(r, n, n_o, n_v, 'left', 'left'), r)
    singles = R.restart_vector(r, n1, n_o, n_v, 'right', 'left')
    assert np.array_equal(singles[:n1], r[:n1]) and not singles[n1:].any()
    with pytest.raises(R.RestartError, match='Did not recognize'):
        R.restart_vector(r, n1 + 1, n_o, n_v, 'right', 'right')


def test_start_vectors_fall_back_to_the_other_side(tmp_path):
    n_o, n_v = 2, 3
    files = R.CCSDFiles(str(tmp_path), n_o, n_v)
    n = files.n_t1 + files.n_t2
    rng = np.random.default_rng(2)
    r1, l2 = rng.normal(size=n), rng.normal(size=n)
    files.right(1).save(r1, 0.3)
    files.left(2).save(l2, 0.4)
    v, e = files.check_and_get_restart_vector(1, 'right')
    assert np.array_equal(v, r1) and e == 0.3
    v, e = files.check_and_get_restart_vector(2, 'right')
    assert np.allclose(v, R.restart_vector(l2, n, n_o, n_v, 'left', 'right')) and e == 0.4
    v, e = files.check_and_get_restart_vector(1, 'left')
    assert np.allclose(v, R.restart_vector(r1, n, n_o, n_v, 'right', 'left'))
    assert files.check_and_get_restart_vector(3, 'left') is None
