"""Known-answer file layout (FK20-CUDA fk20_512_kat.bin, fk20_512test_bootstrap.cu:44-59) and
the falsifiability mangle (fk20_512test.cu:961-1004), on the host."""
import numpy as np
import pytest

from fk20_amd import katfile as K


def test_reference_file_size():
    # fk20_512test.cu:916 bytesExpected = 449970176 at 512 rows, N = 4096
    assert K.file_bytes(512, 4096) == 449970176
    assert K.rows_of(449970176, 4096) == 512


@pytest.mark.parametrize("N", [1024, 2048, 4096])
def test_rows_of_inverts_file_bytes(N):
    for rows in (1, 2, 7, 512):
        assert K.rows_of(K.file_bytes(rows, N), N) == rows
    for bad in (0, K.file_bytes(0, N), K.file_bytes(3, N) + 8):
        with pytest.raises(ValueError):
            K.rows_of(bad, N)


def _arrays(rows, N, seed):
    rng = np.random.default_rng(seed)
    return {name: rng.integers(0, 2 ** 63, size=shp, dtype=np.uint64) for name, shp in K.shapes(rows, N).items()}


def test_write_read_round_trip_and_section_order(tmp_path):
    rows, N = 3, 1024
    arrays = _arrays(rows, N, 1)
    path = str(tmp_path / "kat.bin")
    assert K.write(path, arrays) == K.file_bytes(rows, N)
    back = K.read(path, N)
    for name in K.SECTIONS:
        assert np.array_equal(back[name], arrays[name]), name
    # the sections lie back to back in the reference's order
    raw = np.fromfile(path, dtype=np.uint64)
    off = 0
    for name in K.SECTIONS:
        n = arrays[name].size
        assert np.array_equal(raw[off:off + n], arrays[name].ravel()), name
        off += n
    assert off == raw.size
    with pytest.raises(ValueError):
        K.read(path, N, rows=rows + 1)


def test_write_refuses_existing_file_and_bad_shapes(tmp_path):
    rows, N = 2, 1024
    arrays = _arrays(rows, N, 2)
    path = str(tmp_path / "kat.bin")
    K.write(path, arrays)
    with pytest.raises(FileExistsError):
        K.write(path, arrays)
    arrays2 = _arrays(rows, N, 3)
    K.write(path, arrays2, overwrite=True)
    assert np.array_equal(K.read(path, N)["h_fft"], arrays2["h_fft"])
    bad = dict(arrays)
    bad["h"] = bad["h"][:, :-1]
    with pytest.raises(ValueError):
        K.write(str(tmp_path / "bad.bin"), bad)


def _mangle_loop(a, step, start=3):
    """Element-by-element restatement of varMangle (fk20_512test.cu:971-983)."""
    a = [list(x) for x in np.asarray(a).reshape(-1, np.asarray(a).shape[-1])]
    size = len(a)
    i = start
    while i < size:
        if i + step < size:
            a[i], a[i + 1] = a[i + 1], a[i]
        i += step
    return np.array(a, dtype=np.uint64)


@pytest.mark.parametrize("n,width,step", [(40, 4, 8), (64, 18, 32), (513, 4, 512), (600, 4, 512), (5, 4, 8)])
def test_mangle_matches_reference_loop(n, width, step):
    a = np.arange(n * width, dtype=np.uint64).reshape(n, width)
    got = K.mangle(a, step)
    assert np.array_equal(got.reshape(-1, width), _mangle_loop(a, step))
    assert np.array_equal(K.mangle(got, step), a)  # an involution
    assert np.array_equal(a, np.arange(n * width, dtype=np.uint64).reshape(n, width))  # input untouched
    changed = not np.array_equal(got, a)
    assert changed == (n > 3 + step)


def test_mangle_keeps_row_shape():
    a = np.arange(2 * 16 * 8 * 4, dtype=np.uint64).reshape(2, 16, 8, 4)
    got = K.mangle(a, 32)
    assert got.shape == a.shape
    assert np.array_equal(got.reshape(-1, 4)[3], a.reshape(-1, 4)[4])


def test_toeplitz_vectors_are_zero_where_the_ntt_test_would_swap():
    """Why the NTT stage's falsifiability run mangles its expected output: slots 3 and 4 of
    every Toeplitz vector (the varMangle pairs at step 512) hold zeros."""
    import oracle as O
    from fk20_amd import synthetic as S
    tc = O.poly2toeplitz_coefficients(S.random_polys(1, 1024, 5)[0])
    assert not tc[:, 3:5].any() and tc[:, 0].any()
