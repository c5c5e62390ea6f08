"""Host-side pieces of the densification workload: the dataset_log.csv columns
(densify_pareto_property_network_32.py:35-82), the Stage-II sample
(prepare_stage2_training_data.py:330-354: keys, dtypes, shapes, block average) and
its file name (262-272).  No GPU."""
import csv

import numpy as np

from maple_amd import densify


def test_dataset_log_columns(tmp_path):
    p = tmp_path / "dataset_log.csv"
    row = {k: "" for k in densify.FIELDNAMES}
    row.update({"seed": 7, "vf_target": 0.25, "stiffness": 0.1, "_iters_run": 3})
    densify.write_dataset_log(str(p), [row])
    densify.write_dataset_log(str(p), [row], append=True)
    rows = list(csv.reader(open(p)))
    assert rows[0] == densify.FIELDNAMES and len(densify.FIELDNAMES) == 46
    assert len(rows) == 3 and rows[1][densify.FIELDNAMES.index("seed")] == "7"


def test_stage2_sample(tmp_path):
    rng = np.random.default_rng(0)
    rho = rng.uniform(0, 1, (64, 64, 64)).astype(np.float32)
    cond = np.arange(7, dtype=np.float64)
    p = tmp_path / "s.npz"
    densify.write_stage2_npz(str(p), cond, 42, 3, rho)
    z = np.load(p)
    assert set(z.files) == {"condition", "seed", "source_index", "rho_target_32"}
    assert z["condition"].dtype == np.float32 and z["condition"].shape == (7,)
    assert z["seed"].dtype == np.int64 and int(z["seed"]) == 42 and int(z["source_index"]) == 3
    t = z["rho_target_32"]
    assert t.dtype == np.float16 and t.shape == (1, 32, 32, 32)
    want = rho.reshape(32, 2, 32, 2, 32, 2).mean(axis=(1, 3, 5))
    assert np.abs(t[0].astype(np.float32) - want).max() <= 1e-3
    assert densify.downsample_block_average(rho[:32, :32, :32], 32).shape == (32, 32, 32)


def test_stage2_sample_name():
    assert densify.stage2_sample_name(5, {"seed": "12"}) == "sample-000005_seed-0000000012.npz"
    assert densify.stage2_sample_name(0, {"seed": 3.0, "weight_idx": 4}) == "weight_idx-4_seed-0000000003.npz"
