"""The Node-API binding (pyr3_amd/napi) a TypeScript host would load.

CPU: the addon loads under the installed Node and turns C-ABI errors into JS
exceptions. GPU: one frame rendered from packed blobs through Node (its
readbacks awaited as promises) matches the same blobs rendered through the
Python facade bit for bit.
"""
import json
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT
from helpers import blobs, genome, sierpinski
from pyr3_amd import pack

NAPI = ROOT / "pyr3_amd" / "napi"
NODE = shutil.which("node")
HEADERS = Path("/usr/include/node/node_api.h")
pytestmark = pytest.mark.skipif(NODE is None or shutil.which("gcc") is None or
                                not HEADERS.exists(),
                                reason="node, gcc or the Node-API headers are absent")


@pytest.fixture(scope="module", autouse=True)
def addon():
    """The addon built from pyr3_napi.c against this tree's C ABI (a few
    hundred ms of gcc; the library itself is not rebuilt)."""
    subprocess.run(["sh", str(NAPI / "build.sh")], check=True, timeout=120)
    return NAPI / "pyr3hip.node"


def _node(*args, **kw):
    return subprocess.run([NODE, *args], capture_output=True, text=True, timeout=120, **kw)


def test_addon_loads_and_exports_stage_interface():
    r = _node("-e", "const m=require(process.argv[1]); console.log(Object.keys(m).join(','))",
              str(NAPI / "pyr3hip.node"))
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().split(",") == ["create", "reset", "iterate", "present",
                                           "readHistogram", "readImage", "readIndexMap", "sync"]


def test_addon_raises_c_abi_errors():
    r = _node("-e", "const m=require(process.argv[1]);"
                    "try { m.create(0, 4, 1, new Float32Array([0,1,0])); } "
                    "catch (e) { console.log(e.message); }", str(NAPI / "pyr3hip.node"))
    assert "pyr3hip error -1" in r.stdout


@pytest.mark.gpu
def test_node_render_matches_python_render(tmp_path):
    from pyr3_amd.renderer import Renderer
    g = genome(sierpinski(), scale=20.0, cx=0.5, cy=0.5)
    W, H, os_ = 40, 30, 1
    xf, ds, u, pal = blobs(g, W, H, iters=200, fuse=10, seed=77, walkers=300)
    u = np.array(u, dtype=np.uint32)
    u[14] = 1   # colour-index capture, read back through readIndexMap
    taps = pack.spatial_kernel(0.5, os_)
    k1, k2 = pack.calibration(g.scale, 300 * 200, 1.0, os_)
    de = pack.de_uniforms(W, H, 9, 0, 0.4, k1, k2, os_)
    viz = pack.viz_uniforms(W, H, k1, k2, g.tonemap, os_, len(taps))
    for name, arr in (("taps.f32", taps), ("palette.f32", pal), ("xforms.f32", xf),
                      ("distrib.u32", ds), ("uniforms.u32", u), ("de.u32", de), ("viz.u32", viz)):
        np.ascontiguousarray(arr).tofile(tmp_path / name)
    (tmp_path / "meta.json").write_text(json.dumps({"width": W, "height": H, "oversample": os_,
                                                     "capture": True}))
    r = _node(str(NAPI / "render_blobs.js"), str(tmp_path))
    assert r.returncode == 0, r.stderr
    hn = np.fromfile(tmp_path / "hist_node.u32", dtype=np.uint32).reshape(H, W, 4)
    imn = np.fromfile(tmp_path / "image_node.u8", dtype=np.uint8).reshape(H, W, 4)
    rp = Renderer(W, H, os_)
    rp.reset(g)
    rp.iterate(g, 77, 300, 200, fuse=10, color=pack.ColorOpts(capture_index=True))
    rp.present(g, 300 * 200)
    hp = rp.read_histogram()
    assert np.array_equal(hn, hp)
    assert np.array_equal(imn, rp.read_image())
    idx = np.fromfile(tmp_path / "idx_node.u32", dtype=np.uint32).reshape(H, W)
    cnt = np.fromfile(tmp_path / "cnt_node.u32", dtype=np.uint32).reshape(H, W)
    assert np.array_equal(idx, rp.read_index()) and idx.sum() > 0
    assert np.array_equal(cnt, hp[..., 3])
