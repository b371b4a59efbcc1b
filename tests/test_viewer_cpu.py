"""CPU test of the results viewer (fwi2d_amd/viewer.py): from a run folder
laid out as the executables write it (parameters.txt, model, SPS/RPS, one
shot's seismogram, RTM section, convergence text, final model), it writes one
PNG per figure and skips the figures whose inputs are missing."""
import os

import numpy as np

from fwi2d_amd import configs as C
from fwi2d_amd import viewer


def test_viewer_writes_every_figure(tmp_path):
    nz, nx, nt = 30, 70, 120
    s = C.Survey("v", nz, nx, 10.0, nt, 1e-3, 30.0, 3, 5, 20, 20.0, 10.0)
    root = str(tmp_path)
    folders = C.ensure_dirs(root)
    model = os.path.join(root, "model.bin")
    v = C.flat_layers(nz, nx, [1500.0, 2500.0, 3500.0])
    v.tofile(model)
    xs = np.array([100.0, 300.0, 500.0], np.float32)
    xr = np.linspace(0.0, 690.0, 5).astype(np.float32)
    sps, rps = os.path.join(root, "sps.txt"), os.path.join(root, "rps.txt")
    C.write_geometry(sps, rps, xs, np.full(3, 20.0), xr, np.full(5, 10.0))
    params = os.path.join(root, "parameters.txt")
    C.write_parameters(params, s, model, sps, rps, folders, max_iteration=4)
    rng = np.random.default_rng(0)
    rng.standard_normal((5, nt)).astype(np.float32).tofile(
        os.path.join(folders["data"], f"seismogram_nt{nt}_nr5_1000us_shot_2.bin"))
    suffix = f"30Hz_{nz}x{nx}_10m.bin"
    rng.standard_normal((nx, nz)).astype(np.float32).tofile(
        os.path.join(folders["seismic"], "RTM_section_" + suffix))
    np.savetxt(os.path.join(folders["residuo"], "convergence_4_iterations.txt"),
               [5.0, 4.0, 3.5, 3.2, 3.0])
    (v * 1.01).astype(np.float32).tofile(os.path.join(folders["models"],
                                                      "final_model_FWI_" + suffix))
    out = os.path.join(root, "figures")
    written = viewer.main([params, "--out", out])
    names = sorted(os.path.basename(w) for w in written)
    assert names == ["image_RTM.png", "model.png", "model_FWI.png", "residuo_FWI.png",
                     "seismogram_shot_2.png"]
    for w in written:
        assert os.path.getsize(w) > 1000
    # a missing output is skipped, not an error
    os.remove(os.path.join(folders["seismic"], "RTM_section_" + suffix))
    written = viewer.main([params, "--out", out])
    assert "image_RTM.png" not in [os.path.basename(w) for w in written]
