"""The C++ drop-in (include/bf/LBM.h) drives the GPU solver like the
reference's demos drive the GL solver; its result equals the FP32 oracle run
of the same scenario."""
import os
import subprocess

import numpy as np
import pytest

from helpers import cavity_ids

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DEMO = os.path.join(HERE, "compat", "lid_cavity_demo")


def test_lid_cavity_through_lbm_shim(orc, tmp_path):
    assert os.path.exists(DEMO), "build() compiles tests/compat/lid_cavity_demo"
    out = tmp_path / "cavity.bin"
    ticks = 400
    print(f"case: LBM.h lid cavity 64x64, {ticks} ticks", flush=True)
    r = subprocess.run([DEMO, str(out), str(ticks), str(tmp_path / "ckpt")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # the unchanged render calls reached the registered renderer
    assert "renders 6 bad 0" in r.stdout
    assert (tmp_path / "ckpt" / f"tick_{ticks}.bfmacro").exists()
    raw = out.read_bytes()
    t = np.frombuffer(raw[:4], np.int32)[0]
    n = 64 * 64
    vd = np.frombuffer(raw[4:4 + 16 * n], np.float32).reshape(n, 4)
    sums = np.frombuffer(raw[4 + 16 * n:], np.float64)
    assert t == ticks
    o = orc.Oracle((64, 64, 1), 0.6, periodic=(0, 0, 1), velocity_set="D2Q9", precision="f32")
    o.set_object(2, velocity=(np.float32(1.6) / np.float32(16.0), 0, 0))
    o.init_fields(np.ones(n, np.float32), np.zeros((n, 3), np.float32), boundary_id=cavity_ids(64, 64))
    o.step(ticks, save=True)
    ref = o.read("velocity_density")
    fluid = cavity_ids(64, 64) == 0
    d = np.abs(vd[fluid].astype(np.float64) - ref[fluid]).max()
    print(f"  max|shim - oracle| = {d:.3e}", flush=True)
    assert d <= 1e-6
    np.testing.assert_allclose(sums[0], o.conservation()[0], rtol=1e-12)


DROPLET = os.path.join(HERE, "compat", "droplet_isosurface_demo")


def test_droplet_isosurface_through_marching_cubes_shim(tmp_path):
    """LBM.h + MarchingCubes.h: the mesh from the device-resident density and the
    one from the host copy both equal the oracle's mesh of that density."""
    import mc_oracle as mc
    assert os.path.exists(DROPLET), "build() compiles tests/compat/droplet_isosurface_demo"
    out = tmp_path / "droplet.bin"
    print("case: LBM.h Shan-Chen droplet 32^3 D3Q19, 100 ticks, MarchingCubes.h", flush=True)
    r = subprocess.run([DROPLET, str(out), "100"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    n = 32
    rho = np.frombuffer(raw[:4 * n ** 3], np.float32).reshape(n, n, n)
    off = 4 * n ** 3
    meshes = []
    for _ in range(2):
        nv, ni = np.frombuffer(raw[off:off + 16], np.int64)
        off += 16
        v = np.frombuffer(raw[off:off + 12 * nv], np.float32).reshape(-1, 3)
        off += 12 * nv
        t = np.frombuffer(raw[off:off + 4 * ni], np.uint32).reshape(-1, 3)
        off += 4 * ni
        meshes.append((v, t))
    ov, ot = mc.marching_cubes(rho, 1.0, (n, n, n))
    for v, t in meshes:
        assert np.array_equal(v.view(np.uint32), ov.view(np.uint32)) and np.array_equal(t, ot)
    assert len(ot) > 500



def test_field_copies_to_device_memory(bfa):
    """bf_copy_field_to_device: the velocity/density and force/temperature
    fields arrive in device memory (a torch buffer) bit-identical to the host
    readback, the boundary field in the reference's R8 texture encoding (0
    fluid, 255 object); a wrong size or a field without a texture is refused."""
    import torch
    X, Y, Z = 40, 24, 12
    cells = X * Y * Z
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    ids = (((x - 12) ** 2 + (y - 12) ** 2 + (z - 6) ** 2 <= 16) | (y == 0) | (y == Y - 1)).astype(np.uint8).ravel()
    print(f"case: field copies to device memory, D3Q19 duct {X}x{Y}x{Z}", flush=True)
    s = bfa.Solver((X, Y, Z), 0.6, periodic=(1, 0, 1), velocity_set="D3Q19")
    s.init_fields(np.ones(cells, np.float32), np.zeros((cells, 3), np.float32), boundary_id=ids)
    s.set_param("constant_force", np.array([2e-5, 0, 0], np.float32))
    s.step(25, save=True)
    dev = torch.cuda.current_device()
    for name in ("velocity_density", "force_temperature"):
        buf = torch.full((cells, 4), float("nan"), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()  # the fill runs on torch's stream, the copy on the solver's
        s.copy_field_to_device(name, buf.data_ptr(), buf.numel() * 4)
        s.synchronize()
        got = buf.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), s.read(name).view(np.uint32)), name
    b = torch.full((cells,), 7, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    s.copy_field_to_device("boundary", b.data_ptr(), cells)
    s.synchronize()
    assert np.array_equal(b.cpu().numpy(), np.where(ids != 0, 255, 0).astype(np.uint8))
    with pytest.raises(bfa.BFError):
        s.copy_field_to_device("boundary", b.data_ptr(), cells - 1)
    with pytest.raises(bfa.BFError):
        s.copy_field_to_device("psi", b.data_ptr(), cells)
    s.close()
