"""plot_field / plot_all_fields (simulation.py:630-660) write PNG files; no GPU."""

import numpy as np

import matplotlib

matplotlib.use("Agg")


def test_plot_all_fields_writes_one_image_per_field(tmp_path):
    from pyphasefield_amd import Simulation
    sim = Simulation(dimensions=[12, 10], save_path=str(tmp_path))
    sim.add_field(np.random.default_rng(0).random((12, 10)), "phi")
    sim.add_field(np.zeros((12, 10)), "c")
    sim.time_step_counter = 7
    assert sim.plot_all_fields() == 0
    folder = tmp_path / "images_step_7"
    assert sorted(p.name for p in folder.iterdir()) == ["c_step_7.png", "phi_step_7.png"]
    assert (folder / "phi_step_7.png").read_bytes()[:8] == b"\x89PNG\r\n\x1a\n"
    assert sim.progress_bar() is None and sim.generate_python_script() is None
