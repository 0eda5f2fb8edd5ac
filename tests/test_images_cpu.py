"""Images mode without a GPU: the oracle's IntegratePosition restatement
(SimulationCPU.cpp:159-171, 224-243) against the sigma it must integrate to,
and the PNG writer (Logger.cpp:278-334) against a restatement of the pixel
rules, decoded with zlib."""
import math
import os
import struct
import zlib

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
E, M, HBAR, C1 = 1.60217646e-19, 5.0 * 9.10938188e-31, 1.05457148e-34, 1.15e-9  # constants.h:23-27


def read_png_rgb(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF
        if typ == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 2
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = []
    stride = w * 3 + 1
    for y in range(h):
        assert raw[y * stride] == 0  # filter None
        rows.append(np.frombuffer(raw[y * stride + 1:(y + 1) * stride], dtype=np.uint8))
    return np.stack(rows).reshape(h, w, 3)


def _u8(x):
    """The reference's (unsigned char) cast of a double on x64: truncate to int32, keep the low byte."""
    if not math.isfinite(x) or not (-2147483649.0 < x < 2147483648.0):
        return 0
    return int(x) & 0xFF


def expected_pixels(lt, xx, xy):
    """Logger::WriteImageSection (Logger.cpp:278-318) for the three panels."""
    dim = lt.shape[0]
    px = np.zeros((dim, 3 * dim, 3), dtype=np.uint8)
    for panel, (vals, colored) in enumerate(((lt, False), (xx, False), (xy, True))):
        mn, mx = float(vals.min()), float(vals.max())
        for j in range(dim):
            for i in range(dim):
                v = float(vals[j, i])
                c = panel * dim + i
                if not colored:
                    k = _u8((v - mn) / mx * 255.0)
                    px[j, c] = (k, k, k)
                elif v < 0:
                    px[j, c] = (0, 0, _u8(v / mn * 255.0))
                else:
                    px[j, c] = (_u8(v / mx * 255.0), 0, 0)
    return px


def test_png_writer_pixels(tmp_path):
    import elasticscattering_amd as es

    rng = np.random.default_rng(3)
    dim = 23
    lt = rng.uniform(0, 2e-12, (dim, dim))
    lt[3, 4] = 0.0
    xx = rng.uniform(-0.3, 1.0, (dim, dim)) * 1e-3  # negatives: grey values leave [0, 255] and wrap
    xy = rng.uniform(-1, 1, (dim, dim)) * 1e-5
    p = tmp_path / "T0 MF0.png"
    es.log_images(p, lt, xx, xy)
    got = read_png_rgb(p)
    assert got.shape == (dim, 3 * dim, 3)
    assert np.array_equal(got, expected_pixels(lt, xx, xy))
    # a large image spans several stored deflate blocks
    dim = 160
    big = [rng.uniform(0.1, 1.0, (dim, dim)) for _ in range(3)]
    es.log_images(tmp_path / "big.png", *big)
    assert np.array_equal(read_png_rgb(tmp_path / "big.png"), expected_pixels(*big))


@pytest.mark.parametrize("coherent", [0, 1])
def test_oracle_images_integrate_to_sigma(oracle_esm, coherent):
    """Simpson-integrating the position images over positions gives the
    sigma of DeriveTemperature (same terms, another grouping)."""
    x, y = O.read_points(os.path.join(GOLDEN, "test_impurities.dat"))
    ppr, ppq, tau, alpha, speed, B, T = 31, 7, 1.0, 0.3, 1.67834e5, 12.008, 10.0
    og = O.OracleGrid.from_points(oracle_esm, x, y, 1e-7, 1e-7, 2e-9, 10)
    prm = O.sim_params(ppr, ppq, 1e-7, 1e-7, og.cell_size, tau, alpha, 1, speed, 0.0, nthreads=4)
    sig, _, lts = O.compute_sigmas(oracle_esm, og, prm, coherent, B, [T], want_lifetimes=True)
    img = O.derive_images(oracle_esm, og, prm, coherent, B, lts, T, ppr)
    w = np.array([1.0 if k in (0, ppr - 1) else (2.0 if k % 2 == 0 else 4.0) for k in range(ppr)])
    W = np.outer(w, w)
    omega = E * B / M
    tau_t = tau if coherent else HBAR / (1.380649e-23 * T)
    kf = M * speed / HBAR
    sigma_factor = E * E * kf * kf / (2 * math.pi ** 2 * M * C1) * tau_t / (1 + (omega * tau_t) ** 2)
    f = (1.0 / (3.0 * (ppr - 1))) ** 2 * sigma_factor * 1e-8
    for c, ref in ((1, sig[0].xx_comp), (2, sig[0].xy_comp)):
        got = float((img[c] * W).sum()) * f
        assert abs(got - ref) <= 1e-9 * abs(sig[0].xx_comp), (c, got, ref)
    # the lifetime image is the Simpson phi integral of the truncated lifetimes
    lt = np.minimum(lts.reshape(ppr, ppr, 4, ppq), 15 * tau_t)
    wp = np.array([1.0 if k in (0, ppq - 1) else (2.0 if k % 2 == 0 else 4.0) for k in range(ppq)])
    area = 2 * alpha if not coherent else math.pi / 2 - 2 * alpha
    np.testing.assert_allclose(img[0], (lt * wp).sum(axis=(2, 3)) * area / ((ppq - 1) * 3.0), rtol=1e-12, atol=0)
