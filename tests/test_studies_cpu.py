"""Host logic of the precision study: merging one-precision results."""
import json

from akmc import studies


def _part(name, shas, clusters):
    per = [{"clusters": c, "isolated": c // 2, "mean_size": 1.0 + c * 1e-6, "max_size": 5, "hist": [0, c]}
           for c in clusters]
    return {"workload": "w", "rate_error": {"max_rel": 1e-7}, name: {"per_seed": per, "final_sha256": shas}}


def test_combine_counts_identical_lattices_and_compares_seeds(tmp_path):
    a = tmp_path / "fp32.json"
    b = tmp_path / "fp64.json"
    a.write_text(json.dumps(_part("fp32", ["x", "y", "z"], [100, 104, 98])))
    b.write_text(json.dumps(_part("fp64", ["x", "q", "z"], [101, 103, 99])))
    res = studies.combine([str(a), str(b)])
    assert res["identical_final_lattices"] == 2
    assert res["rate_error"]["max_rel"] == 1e-7
    assert set(res["welch_t"]) == {"clusters", "isolated", "mean_size"}
    assert abs(res["welch_t"]["clusters"]) < 1.0
