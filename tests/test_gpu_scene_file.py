"""Scene files on the GPU: a scene read by mrh_scene_file_load renders like
any other scene (segments equal to the oracle's, relMSE 1e-4), and
TracerSystemI::GenerateScene(path) in the C++ host gives the image the
C-ABI render gives for the same file."""
import os
import subprocess

import numpy as np
import pytest

from meturay_amd import native as N

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.join(os.path.dirname(__file__), "golden", "scenes")
HOST = os.path.join(os.path.dirname(N.LIB_PATH), "_build", "tracer_system_host")


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


def test_cpp_host_generate_scene_from_file(ctx, tmp_path):
    assert os.path.exists(HOST), "tracer_system_host not built (python -m meturay_amd.build)"
    out, fout, sout = tmp_path / "img.bin", tmp_path / "filtered.bin", tmp_path / "scene.bin"
    r = subprocess.run([HOST, N.LIB_PATH, str(out), "64", "48", "1", str(fout), os.path.join(HERE, "room.json"),
                        str(sout)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
    got = np.fromfile(sout, np.float32).reshape(48, 64, 3)
    sc = N.load_scene_file(os.path.join(HERE, "room.json"))
    scene = N.Scene(ctx, sc)
    img = torch.zeros(48 * 64 * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    scene.render(img.data_ptr(), 64, 48, 1, 8, 3, 0, camera=sc["cameras"][0])
    assert got.mean() > 0.005
    assert got.tobytes() == img.cpu().numpy().tobytes()
    scene.close()
