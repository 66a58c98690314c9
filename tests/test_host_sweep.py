"""The sweep writes job.sh's CSV format (header and number formats), so a
row built from an NPZ with the published values reads back as the
published row of resultsRE2D/benchmark_results_cpu_Test1_WA.csv."""
import numpy as np

from jax_wspm_amd.sweep import HEADER, csv_row


def test_header_matches_job_sh():
    assert HEADER.split(",")[:6] == ["test_case", "mesh_size", "solver", "preconditioner", "simulation_time",
                                     "iterations"]
    assert len(HEADER.split(",")) == 13


def test_row_format(tmp_path):
    # values of the published row Test1,25,cg,none (simulation time and memory are machine-dependent)
    f = tmp_path / "numerical_results_Test1.npz"
    np.savez(f, simulation_time=17.57591, iterations=np.array([4, 3] * 26 + [1]), l2_pressure=0.176131,
             linf_pressure=0.974471, l2_saturation=5.03111e-3, linf_saturation=4.13681e-2,
             l2_relative_pressure=1.19571e-2, l2_relative_saturation=3.28611e-2, memory_usage=315.4812)
    row = csv_row(f, "Test1", "25", "cg", "none")
    assert row == ("Test1,25,cg,none,17.5759,183,1.7613e-01,9.7447e-01,5.0311e-03,4.1368e-02,1.1957e-02,"
                   "3.2861e-02,315.48")
