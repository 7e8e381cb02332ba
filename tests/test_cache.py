"""The charge-density solution cache (kassiopeia_amd.cache over kem_cache.cpp; host code only).

Pins:
* MD5: the RFC 1321 test suite (appendix A.5) and Python's hashlib on random lengths across
  the 56 / 64-byte padding boundaries;
* KMD5HashGenerator's double rule (KMD5HashGenerator.hh:117-176): with no masked bits the
  hash is MD5 of the raw bytes; with m masked bits it equals MD5 of the double whose 52-bit
  mantissa is rounded half up (in magnitude) at bit m and truncated there, computed here on the
  integer bit pattern (a different route: no floating-point add, carries into the exponent
  field); values below the threshold are left out of the stream; invariances under
  perturbations below the mask;
* the container stream (KSurfaceContainer.hh:259-264, 495-539; KSurface.hh:164-173): a byte
  layout built here with struct/numpy for a mixed container (arrays of triangles with Dirichlet
  values, rectangles with Neumann fluxes, line segments), every omit combination;
* the store: a save / find round trip returns the charge densities bit for bit; the smallest
  stored threshold at most the tolerance wins; a larger tolerance than stored is needed; other
  voltages on the same shapes are superposed; the .kbd file parsed here (keys with names, class
  names, hashes, labels, object locations and sizes, then the objects) holds the container's
  records and densities and the threshold record."""
import hashlib
import os
import struct

import numpy as np
import pytest

from kassiopeia_amd import cache as C
from kassiopeia_amd import geometry as G


@pytest.fixture(scope="module", autouse=True)
def built():
    from kassiopeia_amd import _lib

    if not os.path.exists(_lib.LIBPATH):
        from kassiopeia_amd import build

        build.build()


# ------------------------------------------------------------------ MD5
RFC1321 = [
    (b"", "d41d8cd98f00b204e9800998ecf8427e"),
    (b"a", "0cc175b9c0f1b6a831c399e269772661"),
    (b"abc", "900150983cd24fb0d6963f7d28e17f72"),
    (b"message digest", "f96b697d7cb7938d525a2f31aaf161d0"),
    (b"abcdefghijklmnopqrstuvwxyz", "c3fcd3d76192e4007dfb496cca67e13b"),
    (b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789", "d174ab98d277d9f5a5611c2c9f419d9f"),
    (b"1234567890" * 8, "57edf4a22be3c955ac49da2e2107b67a"),
]


@pytest.mark.parametrize("msg,digest", RFC1321)
def test_md5_rfc1321_suite(msg, digest):
    assert C.md5_hex(msg) == digest


def test_md5_against_hashlib():
    rng = np.random.default_rng(1321)
    for n in list(range(0, 130)) + [1000, 4096, 65537]:
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert C.md5_hex(data) == hashlib.md5(data).hexdigest(), n


# ------------------------------------------------------------------ the double rule
def masked_bytes(x, m):
    """The bytes KMD5HashGenerator streams for double x with m masked bits, from the integer
    bit pattern: magnitude + 2^(m-1) (round half up), low m bits cleared."""
    bits = np.asarray(x, np.float64).view(np.uint64).astype(object)
    out = []
    for b in np.atleast_1d(bits):
        b = int(b)
        if m > 0:
            sign = b & (1 << 63)
            mag = b & ((1 << 63) - 1)
            mag = (mag + (1 << (m - 1))) & ~((1 << m) - 1)
            b = sign | mag
        out.append(struct.pack("<Q", b))
    return b"".join(out)


def test_no_mask_is_raw_md5():
    x = np.random.default_rng(2).normal(size=50) * 1e3
    assert C.hash_doubles(x, 0, 0.0) == hashlib.md5(x.tobytes()).hexdigest()


@pytest.mark.parametrize("m", [1, 7, 8, 9, 20, 33, 48])
def test_masked_bits_round_half_up(m):
    rng = np.random.default_rng(100 + m)
    x = rng.normal(size=400) * 10.0 ** rng.integers(-8, 9, 400)
    # mantissas with the masked field all ones (a carry into the exponent), exactly half, just below
    one = np.float64(1.5)
    b = int(one.view(np.uint64))
    edge = [b | ((1 << m) - 1), b | (1 << (m - 1)), b | ((1 << (m - 1)) - 1),
            int(np.float64(2.0).view(np.uint64)) - 1]
    edge = np.array(edge, dtype=np.uint64).view(np.float64)
    x = np.concatenate([x, edge, -edge])
    for v in x:
        assert C.hash_doubles([v], m, 0.0) == hashlib.md5(masked_bytes(v, m)).hexdigest(), (m, v)
    assert C.hash_doubles(x, m, 0.0) == hashlib.md5(masked_bytes(x, m)).hexdigest()


def test_perturbation_below_the_mask_is_invisible():
    rng = np.random.default_rng(7)
    x = rng.normal(size=200)
    bits = x.view(np.uint64)
    # keep bit 19 (the rounding bit of 20 masked bits) and change bits 0..18
    low = rng.integers(0, 1 << 19, 200, dtype=np.uint64)
    y = ((bits & ~np.uint64((1 << 19) - 1)) | low).view(np.float64)
    assert C.hash_doubles(x, 20, 1e-14) == C.hash_doubles(y, 20, 1e-14)
    # a change above the mask is seen
    z = x.copy()
    z[17] *= 1 + 1e-6
    assert C.hash_doubles(x, 20, 1e-14) != C.hash_doubles(z, 20, 1e-14)


def test_threshold_leaves_values_out():
    a, b = 1.25, -3.5
    for tiny in (0.0, 1e-15, -9e-15):
        assert C.hash_doubles([a, tiny, b], 20, 1e-14) == C.hash_doubles([a, b], 20, 1e-14)
    assert C.hash_doubles([a, 2e-14, b], 20, 1e-14) != C.hash_doubles([a, b], 20, 1e-14)


# ------------------------------------------------------------------ the container stream
def mixed_mesh():
    cube = G.cube_mesh(2)  # 24 triangles, Dirichlet 1 V
    rect = G.cube_rect_mesh(1)  # 6 rectangles
    rect.bc_type[:] = G.NEUMANN
    rect.bc_flux[:] = np.linspace(2.0, 3.0, rect.n)
    lines = G._mesh(G.line_records(np.array([[0.0, 0, 0], [0, 0, 1]]), np.array([[0.0, 0, 1], [0, 0, 2]]), 1e-3),
                    G.LINESEGMENT, 5.0, 3)
    m = G.concat(cube, rect, lines)
    m.bc_value[:cube.n] = np.linspace(-1.0, 1.0, cube.n)
    return m


def stream_bytes(mesh, sigma, omit, m, thr):
    """The container as KMD5HashGenerator streams it, built from the layout: number of arrays;
    per array its size and (basis, boundary, shape) ids as unsigned shorts; per surface the
    shape's doubles, the boundary value (Dirichlet value / Neumann flux) and the density."""
    nd = {0: 11, 1: 11, 2: 7, 3: 6, 4: 3}

    def dbl(v):
        return b"" if abs(v) < thr else masked_bytes(np.float64(v), m)

    runs = []
    i = 0
    while i < mesh.n:
        j = i + 1
        while j < mesh.n and mesh.shapes[j] == mesh.shapes[i] and mesh.bc_type[j] == mesh.bc_type[i]:
            j += 1
        runs.append((i, j))
        i = j
    out = struct.pack("<I", len(runs))
    for i, j in runs:
        out += struct.pack("<IHHH", j - i, 0, int(mesh.bc_type[i]), int(mesh.shapes[i]))
        for k in range(i, j):
            for v in mesh.recs[k, :nd[int(mesh.shapes[k])]]:
                out += dbl(v)
            if mesh.bc_type[k] == G.DIRICHLET:
                if not omit & C.OMIT_DIRICHLET:
                    out += dbl(mesh.bc_value[k])
            else:
                out += dbl(mesh.bc_flux[k])
            if not omit & C.OMIT_BASIS:
                out += dbl(sigma[k])
    return out


@pytest.mark.parametrize("omit", [0, 1, 2, 3])
@pytest.mark.parametrize("m,thr", [(20, 1e-14), (1, 1e-12), (0, 0.0)])
def test_container_stream_layout(omit, m, thr):
    mesh = mixed_mesh()
    sigma = np.random.default_rng(3).normal(size=mesh.n) * 1e-9
    sigma[3] = 0.0  # below the threshold: left out
    want = hashlib.md5(stream_bytes(mesh, sigma, omit, m, thr)).hexdigest()
    assert C.hash_surfaces(mesh, sigma, omit, m, thr) == want


def test_container_hash_invariances():
    mesh = mixed_mesh()
    s1 = np.random.default_rng(4).normal(size=mesh.n)
    s2 = s1 * 3.0
    shape = C.OMIT_BASIS | C.OMIT_DIRICHLET
    # densities do not enter the shape hashes; Dirichlet values only the shape + boundary hash
    assert C.hash_surfaces(mesh, s1, C.OMIT_BASIS, 20, 1e-14) == C.hash_surfaces(mesh, s2, C.OMIT_BASIS, 20, 1e-14)
    m2 = mesh.with_voltages(np.arange(8) * 10.0)
    assert C.hash_surfaces(mesh, None, shape, 20, 1e-14) == C.hash_surfaces(m2, None, shape, 20, 1e-14)
    assert C.hash_surfaces(mesh, None, C.OMIT_BASIS, 20, 1e-14) != C.hash_surfaces(m2, None, C.OMIT_BASIS, 20, 1e-14)
    # Neumann fluxes are part of the shape hash (only the Dirichlet policy is omitted)
    m3 = mixed_mesh()
    m3.bc_flux[mesh.bc_type == G.NEUMANN] += 1.0
    assert C.hash_surfaces(mesh, None, shape, 20, 1e-14) != C.hash_surfaces(m3, None, shape, 20, 1e-14)
    # geometry below the 20-bit mask is the same shape; above it is not
    m4 = mixed_mesh()
    m4.recs[:, 2:5] *= 1 + 2e-13
    m5 = mixed_mesh()
    m5.recs[5, 2] += 1e-6
    assert C.hash_surfaces(m5, None, shape, 20, 1e-14) != C.hash_surfaces(mesh, None, shape, 20, 1e-14)
    # a relative change of 2e-13 stays below the mask for most values; all must round the same
    # way for the hash to agree, so compare value by value
    same = [masked_bytes(a, 20) == masked_bytes(b, 20) for a, b in zip(mesh.recs.ravel(), m4.recs.ravel())]
    if all(same):
        assert C.hash_surfaces(m4, None, shape, 20, 1e-14) == C.hash_surfaces(mesh, None, shape, 20, 1e-14)


# ------------------------------------------------------------------ the store
def parse_kbd(path):
    """The keys and objects of a .kbd file: key = name, class, hash (strings with a 32-bit
    length), 32-bit label count, labels, 64-bit object location and size; the object follows."""
    data = open(path, "rb").read()
    at, keys = 0, []

    def string(p):
        n = struct.unpack_from("<I", data, p)[0]
        return data[p + 4:p + 4 + n].decode(), p + 4 + n

    while at < len(data):
        name, p = string(at)
        cls, p = string(p)
        h, p = string(p)
        nl = struct.unpack_from("<I", data, p)[0]
        p += 4
        labels = []
        for _ in range(nl):
            lab, p = string(p)
            labels.append(lab)
        loc, size = struct.unpack_from("<QQ", data, p)
        p += 16
        assert loc == p  # the object follows its key
        keys.append(dict(name=name, cls=cls, hash=h, labels=labels, obj=data[loc:loc + size]))
        at = loc + size
    return keys


def test_save_find_roundtrip_and_file_layout(tmp_path):
    mesh = mixed_mesh()
    sigma = np.random.default_rng(5).normal(size=mesh.n) * 1e-8
    d = str(tmp_path / "cache")
    C.save_solution(d, "run.kbd", mesh, sigma, 3e-11)
    kind, got, fname = C.find_solution(d, mesh, 1e-10)
    assert kind == C.FOUND and fname.endswith("run.kbd")
    assert got.tobytes() == sigma.tobytes()  # bit for bit
    # a tolerance below the stored threshold does not take it
    assert C.find_solution(d, mesh, 1e-11)[0] == C.NOT_FOUND
    # the file: a threshold record and the container, with the reference's labels
    keys = parse_kbd(os.path.join(d, "run.kbd"))
    shape_h = C.hash_surfaces(mesh, None, C.OMIT_BASIS | C.OMIT_DIRICHLET, 20, 1e-14)
    bnd_h = C.hash_surfaces(mesh, None, C.OMIT_BASIS, 20, 1e-14)
    sol_h = C.hash_surfaces(mesh, sigma, 0, 1, 1e-12)
    assert [k["cls"] for k in keys] == ["ResidualThreshold", "SurfaceContainer"]
    assert keys[0]["name"] == "ResidualThreshold_" + sol_h
    assert keys[0]["labels"] == ["ResidualThreshold", shape_h, bnd_h, sol_h]
    thr, = struct.unpack_from("<d", keys[0]["obj"], 0)
    n = struct.unpack_from("<I", keys[0]["obj"], 8)[0]
    assert thr == 3e-11 and keys[0]["obj"][12:12 + n].decode() == sol_h
    assert keys[1]["name"] == "SurfaceContainer_" + sol_h and keys[1]["hash"] == sol_h
    assert keys[1]["labels"] == ["SurfaceContainer", sol_h]
    assert keys[1]["obj"] == stream_bytes(mesh, sigma, 0, 0, 0.0)
    # saving the same solution again adds nothing (KEMFile::Write skips a duplicate name)
    size = os.path.getsize(os.path.join(d, "run.kbd"))
    C.save_solution(d, "run.kbd", mesh, sigma, 3e-11)
    assert os.path.getsize(os.path.join(d, "run.kbd")) == size


def test_smallest_threshold_wins_across_files(tmp_path):
    mesh = mixed_mesh()
    d = str(tmp_path)
    rng = np.random.default_rng(6)
    s_a, s_b, s_c = (rng.normal(size=mesh.n) for _ in range(3))
    C.save_solution(d, "a.kbd", mesh, s_a, 5e-9)
    C.save_solution(d, "b.kbd", mesh, s_b, 2e-10)
    C.save_solution(d, "a.kbd", mesh, s_c, 7e-10)
    kind, got, fname = C.find_solution(d, mesh, 1e-8)
    assert kind == C.FOUND and np.array_equal(got, s_b) and fname.endswith("b.kbd")
    kind, got, _ = C.find_solution(d, mesh, 1e-10)
    assert kind == C.NOT_FOUND
    # a container of another shape is not found
    other = mixed_mesh()
    other.recs[0, 2] += 1e-3
    assert C.find_solution(d, other, 1.0)[0] == C.NOT_FOUND


def test_superposition_of_stored_voltage_sets(tmp_path):
    """Two stored solutions of one geometry with electrode voltages (1, 0) and (0, 1) compose
    the solution of (3, -2): rows merge where every system has equal values, the reduced system
    [[1, 0], [0, 1]] w = (3, -2) is solved, sigma = 3 sigma_1 - 2 sigma_2."""
    cube = G.cube_mesh(2)
    top = cube.centroids()[:, 2] > 0.49
    cube.group[:] = np.where(top, 1, 0)
    d = str(tmp_path)
    rng = np.random.default_rng(8)
    s1, s2 = rng.normal(size=cube.n), rng.normal(size=cube.n)
    C.save_solution(d, "v.kbd", cube.with_voltages([1.0, 0.0]), s1, 1e-12)
    C.save_solution(d, "v.kbd", cube.with_voltages([0.0, 1.0]), s2, 1e-12)
    kind, got, _ = C.find_solution(d, cube.with_voltages([3.0, -2.0]), 1e-10)
    assert kind == C.SUPERPOSED
    assert np.abs(got - (3.0 * s1 - 2.0 * s2)).max() <= 1e-14 * np.abs(s1).max()
    # the stored voltage set itself is found as stored, not superposed
    kind, got, _ = C.find_solution(d, cube.with_voltages([0.0, 1.0]), 1e-10)
    assert kind == C.FOUND and np.array_equal(got, s2)
    # three distinct values where the stored sets have two: not spanned
    v = cube.bc_value.copy()
    m3 = cube.with_voltages([1.0, 2.0])
    m3.bc_value[0] = 7.0
    assert C.find_solution(d, m3, 1e-10)[0] == C.NOT_FOUND
    # solutions stored with a threshold above the tolerance are not superposed
    assert C.find_solution(d, cube.with_voltages([3.0, -2.0]), 1e-13)[0] == C.NOT_FOUND
    del v


def test_rejects_symmetry_groups_and_bad_shapes(tmp_path):
    m = G.mirror_groups(G.cube_mesh(2), [0, 0, 0], [1, 0, 0])
    with pytest.raises(C.CacheError):
        C.find_solution(str(tmp_path), m, 1.0)
    bad = G.cube_mesh(2)
    bad.shapes[0] = 9
    with pytest.raises(C.CacheError, match="unknown shape"):
        C.hash_surfaces(bad)
