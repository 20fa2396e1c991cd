"""PNG datasets on the host (no GPU): the PNG codec and loadImages (image.cpp:818-965).

* The PNG reader is checked against files written by an independent encoder (PIL) in every mode the
  datasets can hold (RGB, RGBA, palette, 1/8/16-bit grayscale, grayscale + alpha), against an Adam7
  interlaced file with every filter type written by the small encoder below, and the writer's files are
  read back by PIL.
* loadImages: files listed recursively in sorted order; labels assigned in the order colours are met
  (padding colour first, then column by column through the images in file order); depth read as
  millimetres; dummy depth 1000 without depth images; depth filling (a hand-worked known answer); the
  reference's rejections (non-RGB colour or label image, non-grayscale depth, depth > 50000, average
  depth outside [0.01, 10] m, sizes that differ, missing files).
* CIELab colour (VIGRA's RGB2LabFunctor, image.cpp:89-96): parity unpinned (no vectors in the reference);
  checked against published CIE L*a*b* values of white, black and the primaries (D65).
* ConfusionMatrix / calculatePixelAccuracy (predict.cpp:44-199): the reference's known answers
  (hyperopt_test.cpp:108-174) and a hand-worked 3-class case.
"""
import os
import struct
import zlib

import numpy as np
import pytest
from PIL import Image

import curfil_amd as cf


def _pil(arr, mode=None):
    return Image.fromarray(arr, mode) if mode else Image.fromarray(arr)


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P", "I;16", "1"])
def test_png_reader_matches_pil(tmp_path, mode):
    rng = np.random.default_rng(len(mode))
    H, W = 23, 37
    if mode == "RGB":
        img = _pil(rng.integers(0, 256, (H, W, 3)).astype(np.uint8))
    elif mode == "RGBA":
        img = _pil(rng.integers(0, 256, (H, W, 4)).astype(np.uint8), "RGBA")
    elif mode == "L":
        img = _pil(rng.integers(0, 256, (H, W)).astype(np.uint8))
    elif mode == "LA":
        img = _pil(rng.integers(0, 256, (H, W, 2)).astype(np.uint8), "LA")
    elif mode == "P":
        img = _pil(rng.integers(0, 256, (H, W, 3)).astype(np.uint8)).quantize(13)
    elif mode == "I;16":
        img = Image.fromarray(rng.integers(0, 65536, (H, W)).astype(np.uint16))
    else:
        img = _pil((rng.random((H, W)) > 0.5).astype(np.uint8) * 255).convert("1")
    p = tmp_path / "x.png"
    img.save(p, optimize=(mode == "L"))
    got = cf.read_png(p)
    if mode == "P":
        want = np.array(img.convert("RGB"))
    elif mode == "1":
        want = np.array(img).astype(np.uint8)       # 1-bit samples: 0 / 1
    else:
        want = np.array(img)
    assert got.shape == want.shape
    assert np.array_equal(got.astype(np.int64), want.astype(np.int64))
    assert got.dtype == (np.uint16 if mode == "I;16" else np.uint8)


def _adam7_png(rgb):
    """an interlaced 8-bit RGB PNG, rows filtered with filter type (row index mod 5)"""
    H, W, _ = rgb.shape
    passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
    raw = bytearray()
    for x0, y0, dx, dy in passes:
        sub = rgb[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        prev = np.zeros(sub.shape[1] * 3, np.int64)
        for r, row in enumerate(sub):
            cur = row.reshape(-1).astype(np.int64)
            ft = r % 5
            left = np.concatenate([np.zeros(3, np.int64), cur[:-3]])
            upleft = np.concatenate([np.zeros(3, np.int64), prev[:-3]])
            if ft == 0:
                f = cur
            elif ft == 1:
                f = cur - left
            elif ft == 2:
                f = cur - prev
            elif ft == 3:
                f = cur - (left + prev) // 2
            else:
                p = left + prev - upleft
                pa, pb, pc = abs(p - left), abs(p - prev), abs(p - upleft)
                pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
                f = cur - pred
            raw.append(ft)
            raw.extend((f % 256).astype(np.uint8).tobytes())
            prev = cur

    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)
    ihdr = struct.pack(">IIBBBBB", W, H, 8, 2, 0, 0, 1)
    return b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(bytes(raw))) + chunk(b"IEND", b"")


def test_png_reader_interlaced_and_all_filters(tmp_path):
    rgb = np.random.default_rng(3).integers(0, 256, (13, 19, 3)).astype(np.uint8)
    p = tmp_path / "i.png"
    p.write_bytes(_adam7_png(rgb))
    assert np.array_equal(np.array(Image.open(p)), rgb)      # the encoder is right
    assert np.array_equal(cf.read_png(p), rgb)


def test_png_writer_read_by_pil(tmp_path):
    rng = np.random.default_rng(9)
    rgb = rng.integers(0, 256, (31, 17, 3)).astype(np.uint8)
    dep = rng.integers(0, 65536, (31, 17)).astype(np.uint16)
    cf.write_png(tmp_path / "c.png", rgb)
    cf.write_png(tmp_path / "d.png", dep)
    assert np.array_equal(np.array(Image.open(tmp_path / "c.png")), rgb)
    assert np.array_equal(np.array(Image.open(tmp_path / "d.png")).astype(np.uint16), dep)
    with pytest.raises(cf.CurfilError, match="not a PNG"):
        (tmp_path / "bad.png").write_bytes(b"GIF89a...")
        cf.read_png(tmp_path / "bad.png")
    broken = bytearray((tmp_path / "c.png").read_bytes())
    broken[40] ^= 0xFF
    (tmp_path / "crc.png").write_bytes(bytes(broken))
    with pytest.raises(cf.CurfilError, match="CRC|corrupt"):
        cf.read_png(tmp_path / "crc.png")


def _write_triple(folder, name, rgb, depth, gt):
    os.makedirs(folder, exist_ok=True)
    cf.write_png(os.path.join(folder, name + "_colors.png"), rgb)
    if depth is not None:
        cf.write_png(os.path.join(folder, name + "_depth.png"), depth)
    cf.write_png(os.path.join(folder, name + "_ground_truth.png"), gt)


PALETTE = np.array([[0, 0, 0], [200, 10, 10], [10, 200, 10], [10, 10, 200], [250, 250, 0]], np.uint8)


def _scene(rng, H, W, labels):
    rgb = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    depth = rng.integers(500, 5000, (H, W)).astype(np.uint16)
    depth[rng.random((H, W)) < 0.1] = 0
    lab = rng.choice(labels, (H, W))
    return rgb, depth, PALETTE[lab]


@pytest.fixture
def folder(tmp_path):
    rng = np.random.default_rng(4711)
    root = tmp_path / "train"
    scenes = {}
    for name, (H, W, labels, sub) in {"b": (30, 40, [1, 2], ""), "a": (30, 40, [3, 1, 0], ""),
                                      "c": (24, 32, [4, 2], "sub")}.items():
        scenes[name] = _scene(rng, H, W, labels)
        _write_triple(str(root / sub), name, *scenes[name])
    return root, scenes


def _first_seen(gts):
    """colours in the order a column-by-column scan of each image in turn meets them"""
    order = []
    for gt in gts:
        for c in gt.transpose(1, 0, 2).reshape(-1, 3):
            t = tuple(int(v) for v in c)
            if t not in order:
                order.append(t)
    return order


def test_load_images_order_labels_and_depth(folder):
    root, scenes = folder
    ds = cf.load_images(str(root), ignoredColors=["250,250,0"])
    assert [os.path.relpath(f, root) for f in ds.files] == ["a_colors.png", "b_colors.png", "sub/c_colors.png"]
    assert (ds.count, ds.width, ds.height, ds.padding_label) == (3, 40, 30, 0)
    # padding colour first, then in scan order through a, b, c
    expect = [(250, 250, 0)] + [c for c in _first_seen([scenes[k][2] for k in "abc"]) if c != (250, 250, 0)]
    assert [tuple(c) for c in ds.label_map.colors()] == expect
    for i, k in enumerate("abc"):
        rgb, depth, labels = ds.get(i)
        assert np.array_equal(rgb, scenes[k][0]) and np.array_equal(depth, scenes[k][1])
        assert np.array_equal(ds.label_map.colors()[labels], scenes[k][2])
    assert ds.image_size(2) == (32, 24)
    # a second set shares the map: known colours keep their labels
    test = cf.load_images(str(root / "sub"), ignoredColors=["250,250,0"], label_map=ds.label_map)
    assert np.array_equal(test.get(0)[2], ds.get(2)[2])


def test_dummy_depth_and_depth_filling(tmp_path):
    gt = np.repeat(PALETTE[[1, 2, 3, 1]][None], 3, axis=0)
    rgb = np.zeros((3, 4, 3), np.uint8)
    depth = np.array([[0, 0, 700, 0],
                      [600, 0, 0, 0],
                      [0, 0, 0, 0]], np.uint16)
    _write_triple(str(tmp_path), "x", rgb, depth, gt)
    ds = cf.load_images(str(tmp_path), useDepthImages=False)
    assert np.all(ds.get(0)[1] == 1000)
    # fillDepthFromRight, FromLeft, FromTop, FromBottom (image.cpp:279-377), worked by hand:
    # right: [700 700 700 0], [600 0 0 0], [0 0 0 0]; left: [700 700 700 700], [600 600 600 600], [0 0 0 0];
    # top: row 2 takes row 1; bottom: nothing left
    filled = cf.load_images(str(tmp_path), useDepthFilling=True).get(0)[1]
    assert filled.tolist() == [[700, 700, 700, 700], [600, 600, 600, 600], [600, 600, 600, 600]]


def test_loader_rejections(tmp_path):
    rng = np.random.default_rng(1)
    rgb, depth, gt = _scene(rng, 12, 16, [1, 2])

    def case(name, **kw):
        d = tmp_path / name
        parts = dict(rgb=rgb, depth=depth, gt=gt)
        parts.update(kw)
        _write_triple(str(d), "f", parts["rgb"], parts["depth"], parts["gt"])
        return str(d)
    with pytest.raises(cf.CurfilError, match="non-RGB"):
        cf.load_images(case("gray_colors", rgb=rgb[:, :, 0]))
    with pytest.raises(cf.CurfilError, match="non-RGB"):
        cf.load_images(case("gray_labels", gt=gt[:, :, 0]))
    with pytest.raises(cf.CurfilError, match="invalid depth image format"):
        cf.load_images(case("rgb_depth", depth=rgb))
    with pytest.raises(cf.CurfilError, match="illegal depth value"):
        cf.load_images(case("too_deep", depth=np.full_like(depth, 50001)))
    with pytest.raises(cf.CurfilError, match="illegal average depth"):
        cf.load_images(case("too_far", depth=np.full_like(depth, 20000)))
    with pytest.raises(cf.CurfilError, match="illegal average depth"):
        cf.load_images(case("no_depth", depth=np.zeros_like(depth)))
    with pytest.raises(cf.CurfilError, match="differ"):
        cf.load_images(case("small_depth", depth=depth[:-1]))
    with pytest.raises(cf.CurfilError, match="different sizes"):
        cf.load_images(case("small_gt", gt=gt[:, :-2]))
    missing = case("missing")
    os.remove(os.path.join(missing, "f_ground_truth.png"))
    with pytest.raises(cf.CurfilError, match="cannot open"):
        cf.load_images(missing)


def test_parse_color():
    assert cf.parse_color("50,205,50") == (50, 205, 50)
    with pytest.raises(ValueError, match="out of range"):
        cf.parse_color("0,256,0")
    with pytest.raises(ValueError, match="illegal color"):
        cf.parse_color("red")


def test_cielab_against_published_values(tmp_path):
    colors = np.array([[255, 255, 255], [0, 0, 0], [255, 0, 0], [0, 255, 0], [0, 0, 255]], np.uint8)
    rgb = np.repeat(colors[None], 2, axis=0)                       # 2 x 5 image
    _write_triple(str(tmp_path), "lab", rgb, np.full((2, 5), 1500, np.uint16), PALETTE[[1, 1, 2, 2, 3]][None]
                  .repeat(2, axis=0))
    lab = cf.load_images(str(tmp_path), useCIELab=True).color(0)[:, 0, :].T
    # CIE L*a*b* (D65) of sRGB white, black, red, green, blue; linear RGB 0 / 1 equals sRGB 0 / 255
    want = np.array([[100.0, 0.0, 0.0], [0.0, 0.0, 0.0], [53.24, 80.09, 67.20], [87.73, -86.18, 83.18],
                     [32.30, 79.19, -107.86]])
    assert np.abs(lab - want).max() < 0.05
    assert np.abs(lab[0]).max() < 100.0001 and abs(lab[0][1]) < 1e-4 and abs(lab[0][2]) < 1e-4
    plain = cf.load_images(str(tmp_path)).color(0)
    assert np.array_equal(plain.transpose(1, 2, 0), rgb.astype(np.float32))


def test_confusion_matrix_known_answers():
    # hyperopt_test.cpp:108-174
    m = cf.ConfusionMatrix(3)
    m.data[:] = [[1.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.0, 0.25, 0.75]]
    assert m.averageClassAccuracy(True) == 0.75
    m = cf.ConfusionMatrix(3, ignoredLabels=[0])
    m.data[:] = [[1.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.0, 0.25, 0.75]]
    assert m.averageClassAccuracy(True) == 1.0 - (1.0 - (1.0 + 0.5 + 0.75) / 3.0)
    assert m.averageClassAccuracy(False) == 1.0 - (1.0 - (0.5 + 0.75) / 2.0)


def test_pixel_accuracy_hand_worked():
    gt = np.array([[0, 1, 1, 2], [2, 2, 1, 0]], np.uint8)
    pr = np.array([[1, 1, 2, 2], [2, 0, 1, 0]], np.uint8)
    # with void: correct 5 of 8; without label 0: pixels (1,1,2,2,2,1): correct 4 of 6
    assert cf.calculatePixelAccuracy(pr, gt, True, [0]) == 5 / 8
    cm = cf.ConfusionMatrix(3, [0])
    assert cf.calculatePixelAccuracy(pr, gt, False, [0], cm) == 4 / 6
    assert cm.data.tolist() == [[0, 0, 0], [0, 2, 1], [1, 0, 2]]
    cm.normalize()
    assert np.allclose(cm.data, [[0, 0, 0], [0, 2 / 3, 1 / 3], [1 / 3, 0, 2 / 3]])
    assert abs(cm.averageClassAccuracy(False) - 2 / 3) < 1e-15
    assert abs(cm.averageClassAccuracy(True) - 4 / 9) < 1e-15
    with pytest.raises(RuntimeError, match="already normalized"):
        cm.normalize()


def test_tools_parse_options_like_the_reference(capfd):
    assert cf.run_tool("train", ["--help"]) == 1
    out = capfd.readouterr().out
    for opt in ("folderTraining", "numThresholds", "ignoreColor", "horizontalFlipping", "subsamplingType"):
        assert "--" + opt in out
    assert cf.run_tool("train", ["folder", "3"]) == 1
    assert "'--samplesPerImage' is required" in capfd.readouterr().err
    assert cf.run_tool("predict", ["out", "test"]) == 1
    assert "'--treeFile' is required" in capfd.readouterr().err
    assert cf.run_tool("predict", ["out", "test", "t.json.gz", "--histogramBias", "1.5"]) == 1
    assert "illegal histogram bias" in capfd.readouterr().err
    assert cf.run_tool("train", ["f", "1", "1", "1", "1", "1", "1", "1", "1", "--mode", "cpu"]) == 1
    assert "not available" in capfd.readouterr().err


REFDATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_testdata")


def test_reference_testdata_pngs_decode_like_pil():
    """the reference's own test images (src/testdata: 640x480 colour, 16-bit depth, ground truth)"""
    names = sorted(os.listdir(REFDATA))
    assert len(names) == 12
    for n in names:
        got = cf.read_png(os.path.join(REFDATA, n))
        want = np.array(Image.open(os.path.join(REFDATA, n)))
        assert np.array_equal(got.astype(np.int64), want.astype(np.int64)), n


def test_reference_testdata_loads_as_the_reference_does():
    ds = cf.load_images(REFDATA, useCIELab=True)
    assert ds.count == 4 and (ds.width, ds.height) == (640, 480)
    assert [os.path.basename(f) for f in ds.files] == ["testing1_colors.png", "training1_colors.png",
                                                       "training2_colors.png", "training3_colors.png"]
    # black (the padding colour) first, then the other four ground-truth colours as met
    cols = [tuple(c) for c in ds.label_map.colors()]
    assert cols[0] == (0, 0, 0) and sorted(cols) == [(0, 0, 0), (0, 0, 255), (50, 205, 50), (255, 0, 0),
                                                     (255, 255, 255)]
    rgb, depth, labels = ds.get(1)
    assert depth[depth > 0].min() >= 1455 and depth.max() <= 9489
    lab = ds.color(1)
    assert np.isfinite(lab).all() and lab[0].min() >= 0.0 and lab[0].max() <= 100.0001


def test_depth_filling_reference_known_answer(tmp_path):
    # image_test.cpp:425-456: two valid pixels, (20, 30) at 1.5 m and (30, 20) at 2.5 m, in a 200 x 100
    # image; after filling every row above 30 holds 2.5 m and the others 1.5 m
    depth = np.zeros((100, 200), np.uint16)
    depth[30, 20] = 1500
    depth[20, 30] = 2500
    _write_triple(str(tmp_path), "f", np.zeros((100, 200, 3), np.uint8), depth, np.zeros((100, 200, 3), np.uint8))
    filled = cf.load_images(str(tmp_path), useDepthFilling=True).get(0)[1]
    assert (filled[:30] == 2500).all() and (filled[30:] == 1500).all()
