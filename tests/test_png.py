"""CPU: the library's PNG writer (hh_write_png), read back by a decoder
written from the PNG specification (tests/pngcheck.py). The decoder itself is
pinned by a hand-assembled file: stored rows, filter 0, known pixels."""
import struct
import zlib

import numpy as np
import pytest

import pngcheck


def _hand_png(img):
    h, w, _ = img.shape
    raw = b"".join(b"\x00" + img[r].tobytes() for r in range(h))

    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)

    return (pngcheck.SIGNATURE + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw, 0)) + chunk(b"IEND", b""))


def test_decoder_reads_a_hand_built_file():
    img = np.arange(3 * 5 * 4, dtype=np.uint8).reshape(3, 5, 4)
    got, types = pngcheck.decode(_hand_png(img))
    assert np.array_equal(got, img) and types == [0, 0, 0]


@pytest.mark.parametrize("flip", [False, True])
def test_write_png_round_trip(tmp_path, flip):
    from helios_amd import runtime
    rng = np.random.default_rng(3)
    img = np.zeros((37, 53, 4), np.uint8)
    img[:12] = rng.integers(0, 256, size=(12, 53, 4))                       # noise
    img[12:25] = np.linspace(0, 255, 53).astype(np.uint8)[None, :, None]   # horizontal ramp
    img[25:] = np.linspace(0, 255, 12).astype(np.uint8)[:, None, None]     # vertical ramp
    path = tmp_path / "x.png"
    runtime.write_png(path, img, flip_rows=flip)
    got, types = pngcheck.decode(path.read_bytes())
    assert np.array_equal(got, img[::-1] if flip else img)
    assert len(set(types)) >= 2            # the writer picks filters per row


def test_write_png_errors(tmp_path):
    from helios_amd import runtime
    with pytest.raises(runtime.HeliosError) as e:
        runtime.write_png(tmp_path / "no_such_dir" / "x.png", np.zeros((2, 2, 4), np.uint8))
    assert e.value.status == 7
    with pytest.raises(ValueError):
        runtime.write_png(tmp_path / "y.png", np.zeros((2, 2, 3), np.uint8))
