"""The example harness (gprat_amd/examples/execute.py, the reference's
examples/gprat_cpp/src/execute.cpp) runs unchanged on the reference's data
directory: a core-count sweep over two training sizes writes output.csv with
the reference's columns, one row per measurement."""
import os

import pytest

pytestmark = pytest.mark.gpu


def test_execute_harness_writes_output_csv(tmp_path, golden_dir):
    from gprat_amd.examples import execute
    out = tmp_path / "output.csv"
    rows = execute.run(golden_dir, start=256, end=512, step=2, loop=1, opt_iter=1, n_test=256, n_cores=4, n_tiles=4,
                       n_reg=8, output=str(out))
    lines = out.read_text().splitlines()
    assert lines[0] + "\n" == execute.HEADER
    assert len(lines) == 1 + len(rows) == 1 + 2 * 2  # cores 2 and 4, sizes 256 and 512
    for line in lines[1:]:
        f = line.split(",")
        assert f[0] == "gpu" and int(f[1]) in (2, 4) and int(f[2]) in (256, 512)
        assert all(float(v) > 0 for v in f[7:14])
    # a second run appends without a second header
    execute.run(golden_dir, start=256, end=256, loop=1, n_test=256, n_cores=2, n_tiles=4, output=str(out))
    assert out.read_text().count("Target,Cores") == 1
