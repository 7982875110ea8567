"""The sweep driver: the reference drivers' size ranges and the amt::metric
report (CPU), and a short GPU sweep of every routine."""
import os

import pytest

import helpers


def sweep_mod():
    import importlib.util
    path = os.path.join(os.path.dirname(helpers.pkg().LIB_PATH), "..", "sweep.py")
    spec = importlib.util.spec_from_file_location("omb_sweep", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_reference_ranges():
    """amt::range as the drivers' main() call it."""
    sw = sweep_mod()
    r = sw.reference_range(*sw.SWEEPS["mtm"]["range"])
    assert r[0] == 32 and r[-1] == 3040 and len(r) == 95          # 32, 64, ..., < 3072
    r = sw.reference_range(*sw.SWEEPS["mtv"]["range"])
    assert r == [2.0 ** k for k in range(1, 16)]                  # 2, 4, ..., 32768
    r = sw.reference_range(*sw.SWEEPS["dot"]["range"])
    assert r[:3] == [2, 1026, 2050] and r[-1] < 2 ** 20 and len(r) == 1024
    with pytest.raises(RuntimeError, match="start > end"):
        sw.reference_range(10, 5, 1, False)


def test_metric_report_format():
    """Fields of amt::metric::str() (include/metric.hpp:435-464) and the CSV
    layout of csv() (:478-500)."""
    sw = sweep_mod()
    m = sw.Metric(3, peak=100.0)
    for v in (10.0, 30.0, 20.0):
        m.update("a", v)
    for v in (20.0, 60.0, 40.0):
        m.update("b tensor", v)
    text = m.str("tensor")
    assert text.startswith("Peak Performance: 100 GFlops\n")
    assert "Name: a\n\tMin GFlops: 10\n\tMax GFlops: 30\n" in text
    assert "\tMax SpeedUp with respect to tensor: 2\n\tAvg SpeedUp with respect to tensor: 2\n" in text
    assert "\tMax Peak Utilization in %: 30\n\tAvg GFlops: 20\n\tAvg Peak Utilization in %: 20\n" in text


def test_metric_csv(tmp_path):
    sw = sweep_mod()
    m = sw.Metric(2, peak=1.0)
    m.update("x", 1.5)
    m.update("x", 2.5)
    m.update("y", 3.0)
    m.update("y", 4.0)
    p = tmp_path / "m.csv"
    m.csv(str(p))
    assert p.read_text() == '"x","y"\n1.5,3\n2.5,4\n'


@pytest.mark.gpu
@pytest.mark.parametrize("op,lo,hi,step", [("mtm", 32, 129, 32), ("mtv", 2, 600, 2),
                                           ("dot", 2, 5000, 1024), ("outer", 32, 129, 32),
                                           ("trans", 32, 129, 32)])
def test_short_gpu_sweep(dev, op, lo, hi, step, tmp_path):
    sw = sweep_mod()
    mul = sw.SWEEPS[op]["range"][3]
    sizes = sw.reference_range(lo, hi, step, mul)
    m = sw.run(op, sizes, iters=2, layout="first",
               modes=("plain", "split") if op == "mtm" else ("plain",))
    assert all(len(v["plot"]) == len(sizes) and min(v["plot"]) > 0 for v in m.data.values())
    assert "Peak Performance" in m.str()
    m.csv(str(tmp_path / "s.csv"))
    assert len((tmp_path / "s.csv").read_text().splitlines()) == len(sizes) + 1
