"""Density images (HydroC99_2Domp/Src/Image.c): the oracle against the
reference's own PPM files, and the rules of the image's size, scan and byte.

`tests/golden/images.npz` holds PPM files written by the reference OpenMP
executable (HydroC99_2Domp built in place, see tests/golden/make_golden.py)
for three decks, with the step -> image markers of its step lines.  The
oracle steps the same decks and must reproduce every file byte for byte.
"""
import ctypes as C
import os

import numpy as np
import pytest

import hydro_amd
import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "images.npz"))
DECKS = sorted({k.split("/")[0] for k in GOLD.files})


def ppm(img: np.ndarray) -> bytes:
    h, w, _ = img.shape
    return b"P6\n%d %d\n255\n" % (w, h) + img.tobytes() + b"#EOF\n"


def deck_params(name):
    kv = {}
    for line in bytes(GOLD[f"{name}/deck"]).decode().splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            kv[k.strip()] = v.strip()
    p = oracle.Params(nx=int(kv["nx"]), ny=int(kv["ny"]), dx=float(kv["dx"]), courant_factor=0.8,
                      tend=1e30, testcase=int(kv["testcase"]), nstepmax=int(kv["nstepmax"]))
    return p, int(kv.get("nimage", 0)), float(kv.get("dtimage", 0.0))


def expected_images(name):
    return [bytes(GOLD[k]) for k in sorted(GOLD.files)
            if k.startswith(name + "/IMAGE_")]


def oracle_images(name):
    """Images and step markers of the oracle over the deck, with the cadence
    of HydroC99_2Domp/Src/main.c:212-221, 462-479."""
    p, nimage, dtimage = deck_params(name)
    port = oracle.Port()
    u = port.init(p)
    imgs, marks = [ppm(port.image(u))], []
    state, t, nxt_t = None, 0.0, dtimage
    for step in range(1, p.nstepmax + 1):
        _, state = port.run(p, u, 1, state)
        t = state[0]
        if nimage > 0 and dtimage <= 0:
            if step % nimage == 0:
                marks.append((step, len(imgs)))
                imgs.append(ppm(port.image(u)))
        elif dtimage > 0 and t >= nxt_t:
            nxt_t += dtimage
            marks.append((step, len(imgs)))
            imgs.append(ppm(port.image(u)))
    return imgs, marks


@pytest.mark.parametrize("name", DECKS)
def test_oracle_reproduces_reference_images(name):
    imgs, marks = oracle_images(name)
    ref = expected_images(name)
    assert len(imgs) == len(ref)
    for k, (a, b) in enumerate(zip(imgs, ref)):
        assert a == b, f"{name} image {k} differs"
    assert marks == [tuple(m) for m in GOLD[f"{name}/marks"].tolist()]


def test_image_size_rule():
    """Image.c:124-126: the smallest shrink with nx/shrink, ny/shrink <= 1000."""
    port = oracle.Port()
    lib = port.lib
    w, h = C.c_int(), C.c_int()
    cases = {(1000, 5): (1, 1000, 5), (1001, 5): (2, 500, 2), (3001, 40): (3, 1000, 13),
             (40, 2999): (3, 13, 999), (8192, 8192): (9, 910, 910),
             (23000, 23000): (23, 1000, 1000), (2000, 1): (2, 1000, 0)}
    for (nx, ny), want in cases.items():
        s = lib.ho_image(nx, ny, np.zeros(1), 1000, None, C.byref(w), C.byref(h))
        assert (s, w.value, h.value) == want, (nx, ny)


def test_image_bytes_and_scan():
    """The byte is (uint8_t) fabs(255 * rho / max): truncation below 256, the
    low byte of the int32 conversion above, 0 for NaN / infinite quotients;
    the maximum scan starts at padded cell (0, 0) and skips NaN cells."""
    port = oracle.Port()
    u = np.zeros((4, 5, 10))  # nx = 6, ny = 1
    u[0, 2, 2:8] = [2.0, 1.0, 0.5, -2.4, np.nan, 0.0]
    img = port.image(u)
    assert img.shape == (1, 6, 3)
    assert (img[..., 0] == img[..., 1]).all() and (img[..., 0] == img[..., 2]).all()
    assert img[0, :, 0].tolist() == [255, 127, 63, 50, 0, 0]
    # a NaN where the scan starts makes the maximum NaN: every byte 0
    u[0, 0, 0] = np.nan
    assert (port.image(u) == 0).all()
    # a maximum in a ghost cell counts (the scan covers every padded cell)
    u[0, 0, 0] = 4.0
    assert port.image(u)[0, :, 0].tolist() == [127, 63, 31, 153, 0, 0]


def test_deck_image_keys_are_read(tmp_path, capfd):
    """nimage / dtimage (HydroC99_2Domp/Src/parametres.c:165-166, 202-203)
    are parsed beside hydroparam_t; a deck without them resets both to 0."""
    lib = hydro_amd.lib()
    lib.hydro_amd_image_cadence.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_double)]
    lib.hydro_amd_image_cadence.restype = None
    for k, (text, want) in enumerate([("&RUN\nnimage=7\n/\n", (7, 0.0)),
                                      ("&RUN\n  dtimage = 0.25\n/\n", (0, 0.25)),
                                      ("&RUN\nnstepmax=3\n/\n", (0, 0.0))]):
        d = tmp_path / f"d{k}.nml"
        d.write_text(text)
        hydro_amd.process_args(["-i", str(d)])
        capfd.readouterr()
        n, dt = C.c_int(-1), C.c_double(-1)
        lib.hydro_amd_image_cadence(C.byref(n), C.byref(dt))
        assert (n.value, dt.value) == want, text
