"""Scene files on the GPU: a scene read by mrh_scene_file_load renders like
any other scene (segments equal to the oracle's, relMSE 1e-4), and
TracerSystemI::GenerateScene(path) in the C++ host gives the image the
C-ABI render gives for the same file."""
import os

import numpy as np
import pytest

from meturay_amd import native as N

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.join(os.path.dirname(__file__), "golden", "scenes")


def relmse(a, b, eps=1e-2):
    return float(np.mean((a - b) ** 2 / (b ** 2 + eps)))


@pytest.mark.parametrize("nee", [0, 1])
@pytest.mark.parametrize("cam", [0, 1])
def test_room_matches_oracle(ctx, oracle, nee, cam):
    sc = N.load_scene_file(os.path.join(HERE, "room.json"))
    c = sc["cameras"][cam]
    w, h, S = 64, 48, 2
    scene = N.Scene(ctx, sc)
    img = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = scene.render(img.data_ptr(), w, h, S, 8, 3, 7, camera=c, nee=nee, mis=nee)
    g = img.cpu().numpy().reshape(h, w, 3)
    scene.close()
    ob = oracle.LBVH(sc["tris"], 30)
    ocam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    ref, seg = oracle.render(sc, ob, ocam, w, h, S, 8, 3, 7, threads=8, tree_width=ctx.tune_get("trace_width"),
                             nee=nee, mis=nee)
    assert st.segments == seg
    assert ref.mean() > 0.005
    assert relmse(g, ref) <= 1e-4
